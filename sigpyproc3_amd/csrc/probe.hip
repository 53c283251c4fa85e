// LDS read-rate probe: the peak the sweep kernel's LDS reads are measured against.
//
// Each block of PR_THREADS threads fills PR_WORDS words of LDS, then reads it for `iters`
// rounds in one of two patterns, and folds what it read into one word per thread (written
// out, so no read is dead):
//   pattern 0: ds_read_b128 -- lane l reads the 16 aligned bytes at 16 l of a 1 KiB row;
//   pattern 1: ds_read2st64_b32 -- lane l reads words l and l + 64 (the sweep's window
//              reads: consecutive words per lane, 256-byte row stride).
// Every round reads PR_UNROLL rows at addresses that change with the round, so the loads
// cannot be hoisted. Bytes read per launch: blocks * PR_THREADS * iters * PR_UNROLL * 16
// (pattern 0) or * 8 (pattern 1).
#include "spp_common.hpp"

namespace {

constexpr int PR_THREADS = 512;
constexpr int PR_WORDS = 8192;  // 32 KiB per block: four blocks per CU
constexpr int PR_UNROLL = 8;

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

template <int PATTERN>
__global__ __launch_bounds__(PR_THREADS) void k_probe_lds(int64_t iters, uint32_t* __restrict__ out) {
    __shared__ uint32_t lds[PR_WORDS];
    const int tid = threadIdx.x;
    for (int i = tid; i < PR_WORDS; i += PR_THREADS) lds[i] = static_cast<uint32_t>(i * 2654435761u);
    __syncthreads();
    const int lane = tid & 63;
    const int wave = tid >> 6;
    uint32_t acc = 0;
    for (int64_t it = 0; it < iters; ++it) {
        const int rot = static_cast<int>(it & 7);
#pragma unroll
        for (int u = 0; u < PR_UNROLL; ++u) {
            if constexpr (PATTERN == 0) {
                // row of 256 words: 1 KiB, lane l at word 4 l
                const int row = (wave * PR_UNROLL + u + rot) & (PR_WORDS / 256 - 1);
                const v4u v = *reinterpret_cast<const v4u*>(&lds[row * 256 + 4 * lane]);
                acc ^= v.x + v.y + v.z + v.w;
            } else {
                // row of 128 words (two 64-word halves): lane l reads words l and l + 64
                const int row = (wave * PR_UNROLL + u + rot) & (PR_WORDS / 128 - 1);
                const uint32_t* p = &lds[row * 128 + lane];
                acc ^= p[0] + p[64];
            }
        }
    }
    out[static_cast<int64_t>(blockIdx.x) * PR_THREADS + tid] = acc;
}

}  // namespace

extern "C" int spp_probe_lds(int pattern, int64_t iters, int32_t blocks, uint32_t* out, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(out != nullptr, "null output");
    SPP_REQUIRE(pattern == 0 || pattern == 1, "pattern must be 0 (ds_read_b128) or 1 (ds_read2st64_b32)");
    SPP_REQUIRE(iters > 0 && blocks > 0, "iters and blocks must be > 0");
    hipStream_t st = spp::as_stream(stream);
    if (pattern == 0)
        hipLaunchKernelGGL(k_probe_lds<0>, dim3(blocks), dim3(PR_THREADS), 0, st, iters, out);
    else
        hipLaunchKernelGGL(k_probe_lds<1>, dim3(blocks), dim3(PR_THREADS), 0, st, iters, out);
    SPP_LAUNCHED("k_probe_lds");
    return SPP_OK;
}
