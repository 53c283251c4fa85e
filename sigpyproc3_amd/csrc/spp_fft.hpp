// Template-bank convolution for series whose 5-smooth good size Ng is odd.
//
// kernels.convolve_templates (core/kernels.py:893-905) returns
// irfft(rfft(pad(z)) * rfft(t)) with irfft's default output length 2*(Ng//2) = Ng - 1
// when Ng is odd. That is not the Ng-circular convolution c (the even-Ng result): with
// C = DFT_Ng(c) and Lh = (Ng - 1) / 2 it is
//   y = IDFT_{Ng-1}(Y),  Y[k] = C[k] (0 < k < Lh),  Y[0] = Re C[0],  Y[Lh] = Re C[Lh],
//   Y[Ng-1-k] = conj(Y[k]),
// i.e. a fixed linear map of c. The map is computed here in float64 with two Bluestein
// (chirp-z) transforms over power-of-two FFTs; c itself comes from the direct circular
// sums of the template taps.
#pragma once

#include "spp_common.hpp"

namespace spp {

// One template of a bank: `len` taps at `off` in the tap array, reference bin `ref`, and
// the tap sums sh = sum(h), sh2 = sum(h^2). Over the padded length Ng the zero-mean,
// unit-power normalisation of kernels.normalize_template (core/kernels.py:1008-1027) is
// (h - mu) * inv with mu = sh / Ng and inv = 1 / sqrt(sh2 - sh^2 / Ng) (1 if that is 0):
// see temp_norm().
struct TempDesc {
    int32_t len;
    int32_t ref;
    int64_t off;
    double sh;
    double sh2;
};

// (mu, inv) of a template for the padded length ng, in the same float64 operations on
// the host and the device.
__host__ __device__ inline void temp_norm(const TempDesc& td, int64_t ng, double* mu, double* inv) {
    const double g = static_cast<double>(ng);
    *mu = td.sh / g;
    const double var = td.sh2 - td.sh * td.sh / g;
    *inv = var > 0.0 ? 1.0 / sqrt(var) : 1.0;
}

// pocketfft good_size(n, real=True): the smallest 2^a 3^b 5^c >= n (n itself if n <= 6).
int64_t good_size_real(int64_t n);

// Templates convolved together by one odd_convolve pass for a series of n samples.
int32_t odd_conv_batch(int64_t n, int32_t ntemps);

// Device workspace bytes odd_convolve needs (for min(ntemps, odd_conv_batch) templates).
int64_t odd_conv_workspace(int64_t n, int32_t ntemps);

// convs[i * ldc + t], t < n, for the templates temps[0 .. ntemps) (device descriptors and
// taps, mu/inv computed for Ng). Requires good_size_real(n) odd and greater than n.
// Enqueues on `st` without synchronising; returns an SPP_* status.
int odd_convolve(const float* z, int64_t n, const float* taps, const TempDesc* temps, int32_t ntemps,
                 float* convs, int64_t ldc, void* workspace, int64_t workspace_bytes, hipStream_t st);

}  // namespace spp
