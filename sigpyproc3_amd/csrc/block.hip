// In-memory filterbank blocks (float32, [nchans][nsamps], channel-major): per-row rolls and
// the DM-time transform.
//
// Both read row r at column (t + off_r) mod ncols. The host turns the reference's shifts
// into offsets: a circular roll by s gives off = -s mod ncols (kernels.roll_block,
// core/kernels.py:1059-1099); the valid-columns form gives off = start - s with
// start = max(0, max s) (kernels.roll_block_valid, core/kernels.py:1102-1152), so the
// wrap never triggers there.
//
// The DM-time transform adds the rolled rows of every channel in channel order, starting
// from 0, in float32 (the per-column order of np.sum(..., axis=0) in kernels.dmt_block,
// core/kernels.py:1155-1193), so results are bit-exact to the reference's.
#include "spp_common.hpp"

#include <algorithm>

namespace {

constexpr int RB_THREADS = 256;

__global__ __launch_bounds__(RB_THREADS) void k_roll_rows(const float* __restrict__ in, int64_t ncols,
                                                          const int64_t* __restrict__ off,
                                                          int64_t out_cols, float* __restrict__ out) {
    const int64_t r = blockIdx.y;
    const int64_t o = off[r];
    const float* __restrict__ src = in + r * ncols;
    float* __restrict__ dst = out + r * out_cols;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * RB_THREADS + threadIdx.x; t < out_cols;
         t += static_cast<int64_t>(gridDim.x) * RB_THREADS) {
        int64_t j = t + o;
        if (j >= ncols) j -= ncols;
        dst[t] = src[j];
    }
}

// The same roll with 16-byte stores: a thread writes RV float4s of row blockIdx.y (out rows
// 16-byte aligned: out 16-byte aligned, out_cols % 4 == 0), reading the 4 sources of each as
// plain loads (the source offset is arbitrary; a float4 that straddles the wrap takes its
// elements one by one).
constexpr int RV = 4;
__global__ __launch_bounds__(RB_THREADS) void k_roll_rows4(const float* __restrict__ in, int64_t ncols,
                                                           const int64_t* __restrict__ off,
                                                           int64_t out_cols, float* __restrict__ out) {
    const int64_t r = blockIdx.y;
    const int64_t o = off[r];
    const float* __restrict__ src = in + r * ncols;
    float4* __restrict__ dst = reinterpret_cast<float4*>(out + r * out_cols);
    const int64_t nv = out_cols / 4;
    const int64_t v0 = static_cast<int64_t>(blockIdx.x) * RB_THREADS * RV + threadIdx.x;
    float4 q[RV];
#pragma unroll
    for (int i = 0; i < RV; ++i) {
        const int64_t v = v0 + static_cast<int64_t>(i) * RB_THREADS;
        if (v >= nv) break;
        int64_t j = 4 * v + o;
        if (j >= ncols) j -= ncols;
        if (j + 4 <= ncols) {
            q[i] = make_float4(src[j], src[j + 1], src[j + 2], src[j + 3]);
        } else {
            float e[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) e[k] = src[j + k >= ncols ? j + k - ncols : j + k];
            q[i] = make_float4(e[0], e[1], e[2], e[3]);
        }
    }
#pragma unroll
    for (int i = 0; i < RV; ++i) {
        const int64_t v = v0 + static_cast<int64_t>(i) * RB_THREADS;
        if (v < nv) dst[v] = q[i];
    }
}

// One thread per output sample of one DM row; channels in order.
__global__ __launch_bounds__(RB_THREADS) void k_dmt(const float* __restrict__ in, int64_t nchans,
                                                    int64_t ncols, const int64_t* __restrict__ off,
                                                    int64_t out_cols, float* __restrict__ out) {
    const int64_t i = blockIdx.y;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * RB_THREADS + threadIdx.x;
    if (t >= out_cols) return;
    const int64_t* __restrict__ oi = off + i * nchans;
    float acc = 0.0f;
    for (int64_t c = 0; c < nchans; ++c) {
        int64_t j = t + oi[c];
        if (j >= ncols) j -= ncols;
        acc = acc + in[c * ncols + j];
    }
    out[i * out_cols + t] = acc;
}

int check_offsets(const int64_t* off_host, int64_t n, int64_t ncols) {
    for (int64_t k = 0; k < n; ++k) {
        SPP_REQUIRE(off_host[k] >= 0 && off_host[k] < ncols, "offset %lld outside [0, %lld)",
                    static_cast<long long>(off_host[k]), static_cast<long long>(ncols));
    }
    return SPP_OK;
}

}  // namespace

extern "C" int spp_roll_rows(const float* in, int64_t nrows, int64_t ncols, const int64_t* off_host,
                             const int64_t* off, int64_t out_cols, float* out, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(in != nullptr && off != nullptr && out != nullptr && off_host != nullptr, "null argument");
    SPP_REQUIRE(nrows > 0 && ncols > 0 && out_cols > 0 && out_cols <= ncols,
                "need nrows, ncols > 0 and 0 < out_cols <= ncols");
    SPP_REQUIRE(nrows <= 65535, "at most 65535 rows per call");
    if (check_offsets(off_host, nrows, ncols)) return SPP_EINVAL;
    for (int64_t r = 0; r < nrows; ++r) {
        if (off_host[r] + out_cols > ncols && out_cols < ncols) {
            // only a full-width (circular) output may wrap
            SPP_REQUIRE(false, "row %lld: offset %lld + out_cols %lld wraps in a valid-columns roll",
                        static_cast<long long>(r), static_cast<long long>(off_host[r]),
                        static_cast<long long>(out_cols));
        }
    }
    if (out_cols % 4 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0) {
        const int64_t bx4 = spp::ceil_div(out_cols / 4, int64_t(RB_THREADS) * RV);
        SPP_REQUIRE(bx4 <= 0x7fffffffLL, "too many columns");
        hipLaunchKernelGGL(k_roll_rows4, dim3(static_cast<unsigned>(bx4), static_cast<unsigned>(nrows)),
                           dim3(RB_THREADS), 0, spp::as_stream(stream), in, ncols, off, out_cols, out);
        SPP_LAUNCHED("k_roll_rows4");
        return SPP_OK;
    }
    const int64_t bx = std::min<int64_t>(spp::ceil_div(out_cols, int64_t(RB_THREADS)), 4096);
    hipLaunchKernelGGL(k_roll_rows, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(nrows)),
                       dim3(RB_THREADS), 0, spp::as_stream(stream), in, ncols, off, out_cols, out);
    SPP_LAUNCHED("k_roll_rows");
    return SPP_OK;
}

extern "C" int spp_dmt(const float* in, int64_t nchans, int64_t ncols, const int64_t* off_host,
                       const int64_t* off, int64_t ndm, int64_t out_cols, float* out, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(in != nullptr && off != nullptr && out != nullptr && off_host != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && ncols > 0 && ndm > 0 && out_cols > 0 && out_cols <= ncols,
                "need nchans, ncols, ndm > 0 and 0 < out_cols <= ncols");
    SPP_REQUIRE(ndm <= 65535, "at most 65535 DM rows per call");
    if (check_offsets(off_host, ndm * nchans, ncols)) return SPP_EINVAL;
    if (out_cols < ncols) {
        for (int64_t k = 0; k < ndm * nchans; ++k)
            SPP_REQUIRE(off_host[k] + out_cols <= ncols, "offset %lld + out_cols %lld wraps in a valid-columns transform",
                        static_cast<long long>(off_host[k]), static_cast<long long>(out_cols));
    }
    const int64_t bx = spp::ceil_div(out_cols, int64_t(RB_THREADS));
    SPP_REQUIRE(bx <= 0x7fffffffLL, "too many columns");
    hipLaunchKernelGGL(k_dmt, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(ndm)), dim3(RB_THREADS),
                       0, spp::as_stream(stream), in, nchans, ncols, off, out_cols, out);
    SPP_LAUNCHED("k_dmt");
    return SPP_OK;
}
