// Gulp pre-steps that share the dedispersion path's reads and unpack: per-channel sums
// (bandpass), channel masking on the file's own bytes, and zero-DM removal.
//
// Gulp layout: file order [nsamps][nchans], packed nbits (1/2/4/8, bit order `big`) or
// float32 (nbits == 32). All three are HBM-bound single passes over the gulp.
#include "spp_common.hpp"

#include <cmath>

namespace {

template <int NBITS, bool BIG>
__device__ __forceinline__ float sample_at(const void* in, int64_t row, int64_t row_bytes, int64_t c) {
    if constexpr (NBITS == 32) {
        return static_cast<const float*>(in)[row * (row_bytes / 4) + c];
    } else if constexpr (NBITS == 8) {
        return static_cast<float>(static_cast<const uint8_t*>(in)[row * row_bytes + c]);
    } else {
        constexpr int PER = 8 / NBITS;
        const uint32_t b = static_cast<const uint8_t*>(in)[row * row_bytes + c / PER];
        const int k = static_cast<int>(c % PER);
        const int shift = BIG ? (8 - (k + 1) * NBITS) : k * NBITS;
        return static_cast<float>((b >> shift) & ((1u << NBITS) - 1u));
    }
}

// Bandpass of one gulp, one thread per channel, time in order.
//   SEQ:  out[c] = (...((out[c] + x[0][c]) + x[1][c]) ...) in float32, the order of
//         kernels.extract_bpass (core/kernels.py:277-291);
//   !SEQ: out[c] += float32(sum_t x[t][c] in float64), kernels.extract_bpass_par
//         (core/kernels.py:294-310).
template <int NBITS, bool BIG, bool SEQ>
__global__ __launch_bounds__(256) void k_bandpass(const void* __restrict__ in, int64_t nchans,
                                                  int64_t nsamps, int64_t row_bytes,
                                                  float* __restrict__ out) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (c >= nchans) return;
    if constexpr (SEQ) {
        float acc = out[c];
        int64_t t = 0;
        for (; t + 8 <= nsamps; t += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = sample_at<NBITS, BIG>(in, t + u, row_bytes, c);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + v[u];
        }
        for (; t < nsamps; ++t) acc = acc + sample_at<NBITS, BIG>(in, t, row_bytes, c);
        out[c] = acc;
    } else {
        double acc = 0.0;
        for (int64_t t = 0; t < nsamps; ++t) acc += static_cast<double>(sample_at<NBITS, BIG>(in, t, row_bytes, c));
        out[c] = out[c] + static_cast<float>(acc);
    }
}

// Masking on the file's bytes: byte j of every row becomes (b & keep[j]) | set[j]. The host
// builds keep/set from the mask and value so this equals kernels.mask_channels
// (core/kernels.py:321-345) on the unpacked samples followed by the reference's pack.
__global__ __launch_bounds__(256) void k_mask_bytes(uint8_t* __restrict__ data, int64_t nbytes,
                                                    int64_t row_bytes, const uint8_t* __restrict__ keep,
                                                    const uint8_t* __restrict__ set) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < nbytes; i += stride) {
        const int64_t j = i % row_bytes;
        data[i] = static_cast<uint8_t>((data[i] & keep[j]) | set[j]);
    }
}

// Zero-DM removal (kernels.remove_zerodm, core/kernels.py:469-500): one block per sample row.
//   z = sum_c x[t][c] (exact for integer data; float64 for float32 data),
//   y = fma(-z, w[c], x[t][c]) + b[c] in float64,
//   out = float32(y) for float32 data, else the low 8 bits of y truncated toward zero (the
//   float64 -> uint8 store of the reference on x86-64).
template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_zerodm(const void* __restrict__ in, int64_t nchans,
                                                int64_t row_bytes, const float* __restrict__ w,
                                                const float* __restrict__ bpass, void* __restrict__ out) {
    __shared__ double part[256];
    const int64_t t = blockIdx.x;
    double s = 0.0;
    for (int64_t c = threadIdx.x; c < nchans; c += 256) s += static_cast<double>(sample_at<NBITS, BIG>(in, t, row_bytes, c));
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    const double z = part[0];
    for (int64_t c = threadIdx.x; c < nchans; c += 256) {
        const double x = static_cast<double>(sample_at<NBITS, BIG>(in, t, row_bytes, c));
        const double y = fma(-z, static_cast<double>(w[c]), x) + static_cast<double>(bpass[c]);
        if constexpr (NBITS == 32) {
            static_cast<float*>(out)[t * nchans + c] = static_cast<float>(y);
        } else {
            // In-range guard keeps the conversion defined; |y| < 2^31 for any real bandpass.
            const double yc = fmin(fmax(y, -2147483648.0), 2147483647.0);
            static_cast<uint8_t*>(out)[t * nchans + c] = static_cast<uint8_t>(static_cast<int32_t>(yc) & 0xff);
        }
    }
}

int check_nbits(int nbits) {
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 32,
                "nbits must be 1, 2, 4, 8 or 32, got %d", nbits);
    return SPP_OK;
}

}  // namespace

extern "C" int spp_bandpass(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                            float* out, int parallel, void* stream) {
    spp::clear_error();
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(in != nullptr && out != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0, "nchans must be > 0 and nsamps >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    if (nsamps == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
    const dim3 grid(static_cast<unsigned>(spp::ceil_div(nchans, int64_t(256))));
#define SPP_BP(NB, BG)                                                                                   \
    do {                                                                                                 \
        if (parallel) hipLaunchKernelGGL((k_bandpass<NB, BG, false>), grid, dim3(256), 0, st, in, nchans, \
                                         nsamps, row_bytes, out);                                        \
        else hipLaunchKernelGGL((k_bandpass<NB, BG, true>), grid, dim3(256), 0, st, in, nchans, nsamps,  \
                                row_bytes, out);                                                         \
    } while (0)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_BP(1, false); break;
        case 3: SPP_BP(1, true); break;
        case 4: SPP_BP(2, false); break;
        case 5: SPP_BP(2, true); break;
        case 8: SPP_BP(4, false); break;
        case 9: SPP_BP(4, true); break;
        case 16: case 17: SPP_BP(8, false); break;
        default: SPP_BP(32, false); break;
    }
#undef SPP_BP
    SPP_LAUNCHED("k_bandpass");
    return SPP_OK;
}

extern "C" int spp_mask_bytes(uint8_t* data, int64_t nbytes, int64_t row_bytes, const uint8_t* keep,
                              const uint8_t* set, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(data != nullptr && keep != nullptr && set != nullptr, "null argument");
    SPP_REQUIRE(row_bytes > 0 && nbytes >= 0 && nbytes % row_bytes == 0,
                "nbytes (%lld) must be a multiple of row_bytes (%lld)", static_cast<long long>(nbytes),
                static_cast<long long>(row_bytes));
    if (nbytes == 0) return SPP_OK;
    const int64_t blocks = std::min<int64_t>(spp::ceil_div(nbytes, int64_t(256)), 65536);
    hipLaunchKernelGGL(k_mask_bytes, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       spp::as_stream(stream), data, nbytes, row_bytes, keep, set);
    SPP_LAUNCHED("k_mask_bytes");
    return SPP_OK;
}

extern "C" int spp_remove_zerodm(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                                 const float* chanwts, const float* bpass, void* out, void* stream) {
    spp::clear_error();
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(in != nullptr && chanwts != nullptr && bpass != nullptr && out != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0, "nchans must be > 0 and nsamps >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(nsamps < (int64_t(1) << 31), "nsamps must be < 2^31");
    if (nsamps == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
    const dim3 grid(static_cast<unsigned>(nsamps));
#define SPP_ZD(NB, BG) \
    hipLaunchKernelGGL((k_zerodm<NB, BG>), grid, dim3(256), 0, st, in, nchans, row_bytes, chanwts, bpass, out)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_ZD(1, false); break;
        case 3: SPP_ZD(1, true); break;
        case 4: SPP_ZD(2, false); break;
        case 5: SPP_ZD(2, true); break;
        case 8: SPP_ZD(4, false); break;
        case 9: SPP_ZD(4, true); break;
        case 16: case 17: SPP_ZD(8, false); break;
        default: SPP_ZD(32, false); break;
    }
#undef SPP_ZD
    SPP_LAUNCHED("k_zerodm");
    return SPP_OK;
}
