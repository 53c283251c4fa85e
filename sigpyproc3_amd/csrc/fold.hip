// Folding one gulp into (subintegration, subband, phase bin) cells (kernels.fold,
// core/kernels.py:404-445, as called by Filterbank.fold, base.py:1109-1207).
//
// The host computes each valid sample's cell key (subint, phase bin) with the reference's
// float64 arithmetic and lists the samples of every key in time order. One thread owns one
// (key, band) cell: it continues the cell's float32 sum over its samples in time order and,
// for each sample, over the band's channels in channel order, reading x[s + d_c][c] from the
// packed gulp (a negative row wraps into the gulp, as numba's indexing does; the host keeps
// every |delay| < nsamps and every listed sample < nsamps - max delay). That is the
// reference's accumulation order for the cell, so float32 sums are
// bit-exact; no two threads share a cell, so no atomics are needed.
#include "spp_common.hpp"

namespace {

template <int NBITS, bool BIG>
__device__ __forceinline__ float fold_sample(const void* in, int64_t row, int64_t row_bytes, int64_t c) {
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

template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_fold(const void* __restrict__ in, int64_t nsamps, int64_t row_bytes,
                                              const int32_t* __restrict__ delays,
                                              const int32_t* __restrict__ samples,
                                              const int32_t* __restrict__ key_start,
                                              const int32_t* __restrict__ key_pos, int32_t nkeys,
                                              const int32_t* __restrict__ band_start, int32_t nbands,
                                              int32_t nbins, float* __restrict__ fold,
                                              int32_t* __restrict__ count) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= static_cast<int64_t>(nkeys) * nbands) return;
    const int k = static_cast<int>(idx / nbands);
    const int b = static_cast<int>(idx % nbands);
    const int64_t pos = static_cast<int64_t>(key_pos[k]) + static_cast<int64_t>(b) * nbins;
    const int c0 = band_start[b], c1 = band_start[b + 1];
    float acc = fold[pos];
    int32_t cnt = count[pos];
    for (int i = key_start[k]; i < key_start[k + 1]; ++i) {
        const int64_t s = samples[i];
        for (int c = c0; c < c1; ++c) {
            int64_t row = s + delays[c];
            if (row < 0) row += nsamps;  // numba's wrap of a negative index into the gulp
            acc = acc + fold_sample<NBITS, BIG>(in, row, row_bytes, c);
        }
        cnt += c1 - c0;
    }
    fold[pos] = acc;
    count[pos] = cnt;
}

}  // namespace

extern "C" int spp_fold(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                        const int32_t* delays, const int32_t* samples, int64_t nlisted,
                        const int32_t* key_start, const int32_t* key_pos, int32_t nkeys,
                        const int32_t* band_start, int32_t nbands, int32_t nbins, float* fold,
                        int32_t* count, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 32,
                "nbits must be 1, 2, 4, 8 or 32, got %d", nbits);
    SPP_REQUIRE(in != nullptr && delays != nullptr && fold != nullptr && count != nullptr &&
                    key_start != nullptr && key_pos != nullptr && band_start != nullptr,
                "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps > 0 && nbands > 0 && nbins > 0 && nkeys >= 0 && nlisted >= 0,
                "invalid sizes");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    if (nkeys == 0 || nlisted == 0) return SPP_OK;
    SPP_REQUIRE(samples != nullptr, "null sample list");
    const int64_t threads = static_cast<int64_t>(nkeys) * nbands;
    const dim3 grid(static_cast<unsigned>(spp::ceil_div(threads, int64_t(256))));
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
#define SPP_FOLD(NB, BG)                                                                             \
    hipLaunchKernelGGL((k_fold<NB, BG>), grid, dim3(256), 0, st, in, nsamps, row_bytes, delays, samples, \
                       key_start, key_pos, nkeys, band_start, nbands, nbins, fold, count)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_FOLD(1, false); break;
        case 3: SPP_FOLD(1, true); break;
        case 4: SPP_FOLD(2, false); break;
        case 5: SPP_FOLD(2, true); break;
        case 8: SPP_FOLD(4, false); break;
        case 9: SPP_FOLD(4, true); break;
        case 16: case 17: SPP_FOLD(8, false); break;
        default: SPP_FOLD(32, false); break;
    }
#undef SPP_FOLD
    SPP_LAUNCHED("k_fold");
    return SPP_OK;
}
