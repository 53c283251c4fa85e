// Folding one gulp into (subintegration, subband, phase bin) cells (kernels.fold,
// core/kernels.py:404-445, as called by Filterbank.fold, base.py:1109-1207).
//
// The host computes each valid sample's cell key (subint, phase bin) with the reference's
// float64 arithmetic and lists the samples of every key in time order. A cell (key, band)
// continues its float32 sum over its samples in time order and, for each sample, over the
// band's channels in channel order, reading x[s + d_c][c] from the packed gulp (a negative
// row wraps into the gulp, as numba's indexing does; the host keeps every |delay| < nsamps
// and every listed sample < nsamps - max delay).
//
// Integer data: the per-(sample, band) sums are the gulp's subband sums (the fused gulp
// kernel with a band map, exact integers in float32), and a cell's sum is the integer total of
// its samples' band sums. While the cell's float32 running sum stays an integer within 2^24
// it never rounds, so any order gives the reference's value and the cell is o + T. A cell that
// could leave that range (or float32 data) runs the reference's order sample by sample.
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

// Band of every channel from the band starts (chan_to_sub of the subband pass).
__global__ __launch_bounds__(256) void k_band_map(const int32_t* __restrict__ band_start, int32_t nbands,
                                                  int64_t nchans, int32_t* __restrict__ cts) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (c >= nchans) return;
    int lo = 0, hi = nbands;  // largest b with band_start[b] <= c
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (band_start[mid] <= c) lo = mid;
        else hi = mid;
    }
    cts[c] = lo;
}

// One thread per (key, band) cell, bands fastest: T = sum over the key's samples of the
// band sum sub[s][b] (exact integers). o + T when exact (see the file comment), else the
// reference's order.
template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_fold_gather(const void* __restrict__ in, int64_t nsamps,
                                                     int64_t row_bytes, const int32_t* __restrict__ delays,
                                                     const int32_t* __restrict__ samples,
                                                     const int32_t* __restrict__ key_start,
                                                     const int32_t* __restrict__ key_pos, int32_t nkeys,
                                                     const int32_t* __restrict__ band_start, int32_t nbands,
                                                     int32_t nbins, const float* __restrict__ sub,
                                                     float* __restrict__ fold, int32_t* __restrict__ count) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= static_cast<int64_t>(nkeys) * nbands) return;
    const int k = static_cast<int>(idx / nbands);
    const int b = static_cast<int>(idx % nbands);
    const int64_t pos = static_cast<int64_t>(key_pos[k]) + static_cast<int64_t>(b) * nbins;
    const int i0 = key_start[k], i1 = key_start[k + 1];
    uint64_t tsum = 0;
    int i = i0;
    for (; i + 8 <= i1; i += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = sub[static_cast<int64_t>(samples[i + u]) * nbands + b];
#pragma unroll
        for (int u = 0; u < 8; ++u) tsum += static_cast<uint32_t>(v[u]);
    }
    for (; i < i1; ++i) tsum += static_cast<uint32_t>(sub[static_cast<int64_t>(samples[i]) * nbands + b]);
    const int c0 = band_start[b], c1 = band_start[b + 1];
    const float o = fold[pos];
    constexpr double LIM = 16777216.0;
    const double fin = static_cast<double>(o) + static_cast<double>(tsum);
    if (o == truncf(o) && fabs(static_cast<double>(o)) <= LIM && fabs(fin) <= LIM) {
        fold[pos] = static_cast<float>(fin);
    } else {
        float acc = o;
        for (int j = i0; j < i1; ++j) {
            const int64_t s = samples[j];
            for (int c = c0; c < c1; ++c) {
                int64_t row = s + delays[c];
                if (row < 0) row += nsamps;
                acc = acc + fold_sample<NBITS, BIG>(in, row, row_bytes, c);
            }
        }
        fold[pos] = acc;
    }
    count[pos] += (i1 - i0) * (c1 - c0);
}

constexpr int64_t align256(int64_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" int spp_fold_workspace_size(int nbits, int64_t nchans, int64_t nsamps, int32_t nbands,
                                       int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(bytes != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0 && nbands > 0, "invalid sizes");
    *bytes = (nbits == 32) ? 0 : align256(nsamps * nbands * 4) + align256(nchans * 4);
    return SPP_OK;
}

extern "C" int spp_fold(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                        const int32_t* delays, const int32_t* samples, int64_t nlisted,
                        const int32_t* key_start, const int32_t* key_pos, int32_t nkeys,
                        const int32_t* band_start, int32_t nbands, int32_t nbins, float* fold,
                        int32_t* count, void* workspace, int64_t workspace_bytes, void* stream) {
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
    if (nbits == 32) {
        hipLaunchKernelGGL((k_fold<32, false>), grid, dim3(256), 0, st, in, nsamps, row_bytes, delays, samples,
                           key_start, key_pos, nkeys, band_start, nbands, nbins, fold, count);
        SPP_LAUNCHED("k_fold");
        return SPP_OK;
    }
    int64_t need = 0;
    spp_fold_workspace_size(nbits, nchans, nsamps, nbands, &need);
    SPP_REQUIRE(workspace != nullptr, "null workspace");
    SPP_REQUIRE(workspace_bytes >= need, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(need));
    auto* sub = static_cast<float*>(workspace);
    auto* cts = reinterpret_cast<int32_t*>(static_cast<char*>(workspace) + align256(nsamps * nbands * 4));
    hipLaunchKernelGGL(k_band_map, dim3(static_cast<unsigned>(spp::ceil_div(nchans, int64_t(256)))), dim3(256),
                       0, st, band_start, nbands, nchans, cts);
    SPP_LAUNCHED("k_band_map");
    // Band sums of every row (rows whose delayed samples run past the gulp are never listed).
    const int rc = spp_subband(in, nbits, big, nchans, nsamps, delays, cts, 0, nbands, sub,
                               nsamps * nbands, stream);
    if (rc) return rc;
#define SPP_FOLDG(NB, BG)                                                                              \
    hipLaunchKernelGGL((k_fold_gather<NB, BG>), grid, dim3(256), 0, st, in, nsamps, row_bytes, delays,  \
                       samples, key_start, key_pos, nkeys, band_start, nbands, nbins, sub, fold, count)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_FOLDG(1, false); break;
        case 3: SPP_FOLDG(1, true); break;
        case 4: SPP_FOLDG(2, false); break;
        case 5: SPP_FOLDG(2, true); break;
        case 8: SPP_FOLDG(4, false); break;
        case 9: SPP_FOLDG(4, true); break;
        default: SPP_FOLDG(8, false); break;
    }
#undef SPP_FOLDG
    SPP_LAUNCHED("k_fold_gather");
    return SPP_OK;
}
