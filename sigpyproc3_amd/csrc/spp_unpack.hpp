// Sub-byte sample expansion shared by the gulp kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spp {

// The samples of one packed dword as bytes: sample i of the dword (file order: byte by byte,
// and within a byte the SIGPROC bit order `BIG` gives) becomes byte i % 4 of o[i / 4].
// Shift-or spreads for 1/2-bit fields, a byte permute for 4-bit fields.
template <int NBITS, bool BIG>
__device__ __forceinline__ void expand_samples(uint32_t w, uint32_t (&o)[8 / NBITS]) {
    if constexpr (NBITS == 8) {
        o[0] = w;
    } else if constexpr (NBITS == 4) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t p = __builtin_amdgcn_perm(w, w, h ? 0x03030202u : 0x01010000u);  // b, b, b', b'
            o[h] = BIG ? (((p >> 4) & 0x000f000fu) | (p & 0x0f000f00u))
                       : ((p & 0x000f000fu) | ((p >> 4) & 0x0f000f00u));
        }
    } else if constexpr (NBITS == 2) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t b = (w >> (8 * j)) & 0xffu;
            o[j] = BIG ? (((b << 24) | (b << 14) | (b << 4) | (b >> 6)) & 0x03030303u)
                       : ((b | (b << 6) | (b << 12) | (b << 18)) & 0x03030303u);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t b = (w >> (8 * j)) & 0xffu;
            if constexpr (BIG) {
                o[2 * j] = ((b >> 7) | (b << 2) | (b << 11) | (b << 20)) & 0x01010101u;
                o[2 * j + 1] = ((b >> 3) | (b << 6) | (b << 15) | (b << 24)) & 0x01010101u;
            } else {
                o[2 * j] = (b | (b << 7) | (b << 14) | (b << 21)) & 0x01010101u;
                o[2 * j + 1] = ((b >> 4) | (b << 3) | (b << 10) | (b << 17)) & 0x01010101u;
            }
        }
    }
}

}  // namespace spp
