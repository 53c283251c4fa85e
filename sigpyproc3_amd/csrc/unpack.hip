// Sub-byte unpack: 1/2/4-bit samples packed in bytes -> one byte per sample.
//
// HBM-bound (read n bytes, write n * 8/nbits bytes). Each thread expands one 32-bit
// word of packed input into 4 * 8/nbits output bytes with vector stores; a scalar
// path covers misaligned pointers and the tail.
#include "spp_common.hpp"

namespace {

// Expand one packed byte into 8/NBITS samples, first sample in the lowest output byte.
template <int NBITS, bool BIG>
__device__ __forceinline__ uint64_t expand_byte(uint32_t b) {
    constexpr int PER = 8 / NBITS;
    constexpr uint32_t MASK = (1u << NBITS) - 1u;
    uint64_t r = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int shift = BIG ? (8 - (k + 1) * NBITS) : k * NBITS;
        r |= static_cast<uint64_t>((b >> shift) & MASK) << (8 * k);
    }
    return r;
}

template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_unpack_vec(const uint32_t* __restrict__ in, int64_t nwords,
                                                    uint8_t* __restrict__ out) {
    constexpr int PER = 8 / NBITS;  // samples per byte
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords;
         i += stride) {
        const uint32_t w = __builtin_nontemporal_load(in + i);
        uint8_t* o = out + i * (4 * PER);
        if constexpr (PER == 2) {
            uint32_t lo = static_cast<uint32_t>(expand_byte<NBITS, BIG>(w & 0xff)) |
                          (static_cast<uint32_t>(expand_byte<NBITS, BIG>((w >> 8) & 0xff)) << 16);
            uint32_t hi = static_cast<uint32_t>(expand_byte<NBITS, BIG>((w >> 16) & 0xff)) |
                          (static_cast<uint32_t>(expand_byte<NBITS, BIG>(w >> 24)) << 16);
            *reinterpret_cast<uint2*>(o) = make_uint2(lo, hi);
        } else if constexpr (PER == 4) {
            uint4 v;
            v.x = static_cast<uint32_t>(expand_byte<NBITS, BIG>(w & 0xff));
            v.y = static_cast<uint32_t>(expand_byte<NBITS, BIG>((w >> 8) & 0xff));
            v.z = static_cast<uint32_t>(expand_byte<NBITS, BIG>((w >> 16) & 0xff));
            v.w = static_cast<uint32_t>(expand_byte<NBITS, BIG>(w >> 24));
            *reinterpret_cast<uint4*>(o) = v;
        } else {  // PER == 8
            const uint64_t a = expand_byte<NBITS, BIG>(w & 0xff);
            const uint64_t b = expand_byte<NBITS, BIG>((w >> 8) & 0xff);
            const uint64_t c = expand_byte<NBITS, BIG>((w >> 16) & 0xff);
            const uint64_t d = expand_byte<NBITS, BIG>(w >> 24);
            uint4 v0 = make_uint4(static_cast<uint32_t>(a), static_cast<uint32_t>(a >> 32),
                                  static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32));
            uint4 v1 = make_uint4(static_cast<uint32_t>(c), static_cast<uint32_t>(c >> 32),
                                  static_cast<uint32_t>(d), static_cast<uint32_t>(d >> 32));
            reinterpret_cast<uint4*>(o)[0] = v0;
            reinterpret_cast<uint4*>(o)[1] = v1;
        }
    }
}

template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_unpack_scalar(const uint8_t* __restrict__ in, int64_t begin,
                                                       int64_t nbytes, uint8_t* __restrict__ out) {
    constexpr int PER = 8 / NBITS;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = begin + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nbytes;
         i += stride) {
        const uint64_t e = expand_byte<NBITS, BIG>(in[i]);
#pragma unroll
        for (int k = 0; k < PER; ++k) out[i * PER + k] = static_cast<uint8_t>(e >> (8 * k));
    }
}

template <int NBITS, bool BIG>
int launch_unpack(const uint8_t* in, int64_t nbytes, uint8_t* out, hipStream_t st) {
    constexpr int PER = 8 / NBITS;
    const int cus = spp::device_cu_count();
    const bool aligned = (reinterpret_cast<uintptr_t>(in) % 4 == 0) &&
                         (reinterpret_cast<uintptr_t>(out) % (PER == 2 ? 8 : 16) == 0);
    int64_t done = 0;
    if (aligned) {
        const int64_t nwords = nbytes / 4;
        if (nwords > 0) {
            const int64_t blocks = std::min<int64_t>(spp::ceil_div(nwords, 256), int64_t(cus) * 16);
            hipLaunchKernelGGL((k_unpack_vec<NBITS, BIG>), dim3(blocks), dim3(256), 0, st,
                               reinterpret_cast<const uint32_t*>(in), nwords, out);
            SPP_LAUNCHED("k_unpack_vec");
        }
        done = nwords * 4;
    }
    if (done < nbytes) {
        const int64_t rem = nbytes - done;
        const int64_t blocks = std::min<int64_t>(spp::ceil_div(rem, 256), int64_t(cus) * 16);
        hipLaunchKernelGGL((k_unpack_scalar<NBITS, BIG>), dim3(blocks), dim3(256), 0, st, in, done,
                           nbytes, out);
        SPP_LAUNCHED("k_unpack_scalar");
    }
    return SPP_OK;
}

}  // namespace

extern "C" int spp_unpack(const uint8_t* in, int64_t nbytes, uint8_t* out, int nbits, int big,
                          void* stream) {
    spp::clear_error();
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4, "nbits must be 1, 2, or 4, got %d", nbits);
    SPP_REQUIRE(nbytes >= 0, "nbytes must be >= 0, got %lld", static_cast<long long>(nbytes));
    if (nbytes == 0) return SPP_OK;
    SPP_REQUIRE(in != nullptr && out != nullptr, "null buffer");
    hipStream_t st = spp::as_stream(stream);
    if (nbits == 1) return big ? launch_unpack<1, true>(in, nbytes, out, st) : launch_unpack<1, false>(in, nbytes, out, st);
    if (nbits == 2) return big ? launch_unpack<2, true>(in, nbytes, out, st) : launch_unpack<2, false>(in, nbytes, out, st);
    return big ? launch_unpack<4, true>(in, nbytes, out, st) : launch_unpack<4, false>(in, nbytes, out, st);
}
