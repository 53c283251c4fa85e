// DM-trial sweep: plane[i][t] = sum_c x_c[t + d_i(c)] for every DM row i of a delay
// table, over a whole stream resident in HBM.
//
// Stage 1 (k_build_g) turns the packed time-major stream [nsamps][nchans] into a
// channel-major, time-interleaved word array G[c][j] (uint32):
//     G[c][j] = x_c[j] | x_c[j+S] << 8 | x_c[j+2S] << 16 | x_c[j+3S] << 24,  S = 256,
// one byte per sample, zero past the end. A delay shift d of channel c moves every
// byte of a word together, so an output word (4 samples t, t+S, t+2S, t+3S) is the sum
// over channels of the aligned words G[c][t + d_i(c)].
//
// Stage 2 (k_sweep): a block owns a DM tile (<= 16 consecutive DM rows whose
// per-channel delay spread is <= SW_SPREAD_MAX, formed by the host planner) and 1024
// output samples. For each stage of 16 channels it copies G[c][jb .. jb + 256 + spread]
// into LDS (jb = t0 + min delay of the tile, rounded down to 4 words), then each wave
// adds, for its 4 DMs, 4 LDS words per lane per channel (lanes read consecutive words:
// bank-conflict free). Accumulation is SWAR: for nbits <= 4 the four byte lanes of a
// uint32 accumulate up to floor(255/maxval) channels before they are widened to 16-bit
// lanes; for 8-bit data the even/odd bytes go straight to 16-bit lanes. Integer sums
// are exact, so the float32 output is bit-exact to the reference's float64 sums.
//
// Work: nchans adds per DM-trial sample. The sweep is bound by LDS reads (1 byte per
// sample-add) and VALU, not HBM.
#include "spp_common.hpp"

#include <algorithm>
#include <new>
#include <vector>

namespace {

constexpr int SW_THREADS = 256;
constexpr int SW_WAVES = 4;
constexpr int SW_K = 4;                  // DM rows per wave
constexpr int SW_DT = SW_WAVES * SW_K;   // DM rows per tile
constexpr int SW_S = 256;                // interleave stride (words per wave row)
constexpr int SW_T = 4 * SW_S;           // output samples per block
constexpr int SW_CC = 16;                // channels per LDS stage
constexpr int SW_LW = 512;               // LDS words per staged channel
constexpr int SW_SPREAD_MAX = SW_LW - SW_S - 4;
constexpr int G_BUILD_T = 64;            // words per build block (time)
constexpr int G_BUILD_C = 64;            // channels per build block
constexpr int G_BUILD_STRIDE = 68;       // bytes per staged row in the build kernel

template <int NBITS, bool BIG>
__device__ __forceinline__ uint32_t packed_sample(const uint8_t* __restrict__ in, int64_t row,
                                                  int64_t row_bytes, int64_t c) {
    if constexpr (NBITS == 8) {
        return in[row * row_bytes + c];
    } else {
        constexpr int PER = 8 / NBITS;
        const uint32_t b = in[row * row_bytes + c / PER];
        const int k = static_cast<int>(c % PER);
        const int shift = BIG ? (8 - (k + 1) * NBITS) : k * NBITS;
        return (b >> shift) & ((1u << NBITS) - 1u);
    }
}

// Stage 1. Block (bx, by): words [bx*64, bx*64+64) of channels [by*64, by*64+64).
template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_build_g(const uint8_t* __restrict__ in, int64_t nsamps,
                                                 int64_t nchans, uint32_t* __restrict__ G,
                                                 int64_t gpitch) {
    __shared__ uint8_t smp[4 * G_BUILD_T * G_BUILD_STRIDE];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * G_BUILD_T;
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * G_BUILD_C;
    const int64_t row_bytes = nchans * NBITS / 8;
    const int64_t c = c0 + lane;
    // Load rows j0 + r*S + i (r < 4, i < 64): lane = channel, coalesced along the row.
    for (int rr = wave; rr < 4 * G_BUILD_T; rr += 4) {
        const int r = rr / G_BUILD_T, i = rr % G_BUILD_T;
        const int64_t row = j0 + static_cast<int64_t>(r) * SW_S + i;
        uint32_t v = 0;
        if (row < nsamps && c < nchans) v = packed_sample<NBITS, BIG>(in, row, row_bytes, c);
        smp[(r * G_BUILD_T + i) * G_BUILD_STRIDE + lane] = static_cast<uint8_t>(v);
    }
    __syncthreads();
    // Write G[c0 + k][j0 + lane]: lane = word, coalesced along the channel row.
    const int64_t j = j0 + lane;
    if (j >= gpitch) return;
    for (int k = wave; k < G_BUILD_C; k += 4) {
        if (c0 + k >= nchans) break;
        const int i = lane;
        const uint32_t w = static_cast<uint32_t>(smp[(0 * G_BUILD_T + i) * G_BUILD_STRIDE + k]) |
                           (static_cast<uint32_t>(smp[(1 * G_BUILD_T + i) * G_BUILD_STRIDE + k]) << 8) |
                           (static_cast<uint32_t>(smp[(2 * G_BUILD_T + i) * G_BUILD_STRIDE + k]) << 16) |
                           (static_cast<uint32_t>(smp[(3 * G_BUILD_T + i) * G_BUILD_STRIDE + k]) << 24);
        G[(c0 + k) * gpitch + j] = w;
    }
}

// Stage 2.
template <int NBITS, bool USE32>
__global__ __launch_bounds__(SW_THREADS) void k_sweep(
    const uint32_t* __restrict__ G, int64_t gpitch, int64_t nchans,
    const int32_t* __restrict__ table, const int32_t* __restrict__ base,
    const int32_t* __restrict__ tile_dm0, const int32_t* __restrict__ tile_ndm,
    const int32_t* __restrict__ tile_spread, const int32_t* __restrict__ row_len,
    float* __restrict__ plane, int64_t ld, int64_t ntiles_dm, int64_t nblocks) {
    constexpr uint32_t MAXV = (NBITS == 8) ? 255u : ((1u << NBITS) - 1u);
    constexpr int F8 = static_cast<int>(255u / MAXV);     // channels per byte-lane flush
    constexpr int F16 = static_cast<int>(65535u / MAXV);  // channels per 16-bit-lane flush
    __shared__ uint32_t lds[SW_CC * SW_LW];
    __shared__ int32_t srel[SW_DT * SW_CC];

    // XCD-aware order: consecutive blocks on one XCD walk the DM tiles of one time tile.
    const int64_t b = blockIdx.x;
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t logical = (b % 8) * per_xcd + b / 8;
    if (logical >= nblocks) return;
    const int64_t dt = logical % ntiles_dm;
    const int64_t t0 = (logical / ntiles_dm) * SW_T;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int dm0 = tile_dm0[dt];
    const int ndm_t = tile_ndm[dt];
    const int stage_words = SW_S + tile_spread[dt] + 4;
    const int32_t* __restrict__ tbase = base + dt * nchans;

    uint32_t acc8[SW_K][4];
    uint32_t acc16[SW_K][4][2];
    uint32_t acc32[USE32 ? SW_K : 1][USE32 ? 16 : 1];
#pragma unroll
    for (int k = 0; k < SW_K; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc8[k][q] = 0;
            acc16[k][q][0] = 0;
            acc16[k][q][1] = 0;
        }
    if constexpr (USE32) {
#pragma unroll
        for (int k = 0; k < SW_K; ++k)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc32[k][i] = 0;
    }

    auto flush8 = [&]() {
#pragma unroll
        for (int k = 0; k < SW_K; ++k)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc16[k][q][0] += acc8[k][q] & 0x00ff00ffu;
                acc16[k][q][1] += (acc8[k][q] >> 8) & 0x00ff00ffu;
                acc8[k][q] = 0;
            }
    };
    auto flush16 = [&]() {
        if constexpr (USE32) {
#pragma unroll
            for (int k = 0; k < SW_K; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    acc32[k][q * 4 + 0] += acc16[k][q][0] & 0xffffu;
                    acc32[k][q * 4 + 1] += acc16[k][q][1] & 0xffffu;
                    acc32[k][q * 4 + 2] += acc16[k][q][0] >> 16;
                    acc32[k][q * 4 + 3] += acc16[k][q][1] >> 16;
                    acc16[k][q][0] = 0;
                    acc16[k][q][1] = 0;
                }
        }
    };

    int since8 = 0, since16 = 0;
    for (int64_t c0 = 0; c0 < nchans; c0 += SW_CC) {
        const int ncl = static_cast<int>(min<int64_t>(SW_CC, nchans - c0));
        if constexpr (NBITS < 8) {
            if (since8 + ncl > F8) {
                flush8();
                since16 += since8;
                since8 = 0;
            }
        }
        if constexpr (USE32) {
            if (since16 + since8 + ncl > F16) {
                if constexpr (NBITS < 8) {
                    flush8();
                    since16 += since8;
                    since8 = 0;
                }
                flush16();
                since16 = 0;
            }
        }
        __syncthreads();  // the previous stage's LDS reads are done
        {
            // Per-(DM, channel) LDS word offsets of this stage.
            const int dmi = tid / SW_CC, cl = tid % SW_CC;
            if (cl < ncl) {
                const int64_t c = c0 + cl;
                const int64_t jb4 = (t0 + tbase[c]) & ~int64_t(3);
                const int row = dm0 + min(dmi, ndm_t - 1);
                srel[tid] = static_cast<int32_t>(t0 - jb4) + table[static_cast<int64_t>(row) * nchans + c] +
                            cl * SW_LW;
            }
            // Stage G[c][jb4 .. jb4 + stage_words) of this stage's channels.
            for (int s = wave; s < ncl; s += SW_WAVES) {
                const int64_t c = c0 + s;
                const int64_t jb4 = (t0 + tbase[c]) & ~int64_t(3);
                const uint32_t* __restrict__ src = G + c * gpitch + jb4;
                for (int m = lane * 4; m < stage_words; m += 256) {
                    uint4 v = make_uint4(0u, 0u, 0u, 0u);
                    if (jb4 + m < gpitch) v = *reinterpret_cast<const uint4*>(src + m);
                    *reinterpret_cast<uint4*>(&lds[s * SW_LW + m]) = v;
                }
            }
        }
        __syncthreads();
        for (int cl = 0; cl < ncl; ++cl) {
#pragma unroll
            for (int k = 0; k < SW_K; ++k) {
                const int rel = srel[(wave * SW_K + k) * SW_CC + cl];
                const uint32_t* p = lds + rel + lane;
                const uint32_t w0 = p[0], w1 = p[64], w2 = p[128], w3 = p[192];
                if constexpr (NBITS < 8) {
                    acc8[k][0] += w0;
                    acc8[k][1] += w1;
                    acc8[k][2] += w2;
                    acc8[k][3] += w3;
                } else {
                    acc16[k][0][0] += w0 & 0x00ff00ffu;
                    acc16[k][0][1] += (w0 >> 8) & 0x00ff00ffu;
                    acc16[k][1][0] += w1 & 0x00ff00ffu;
                    acc16[k][1][1] += (w1 >> 8) & 0x00ff00ffu;
                    acc16[k][2][0] += w2 & 0x00ff00ffu;
                    acc16[k][2][1] += (w2 >> 8) & 0x00ff00ffu;
                    acc16[k][3][0] += w3 & 0x00ff00ffu;
                    acc16[k][3][1] += (w3 >> 8) & 0x00ff00ffu;
                }
            }
        }
        if constexpr (NBITS < 8) since8 += ncl; else since16 += ncl;
    }
    if constexpr (NBITS < 8) flush8();

    // Store: sample (k, q, r) is t0 + lane + 64q + 256r of DM row dm0 + wave*SW_K + k.
#pragma unroll
    for (int k = 0; k < SW_K; ++k) {
        const int idx = wave * SW_K + k;
        if (idx >= ndm_t) break;
        const int64_t row = dm0 + idx;
        const int64_t len = row_len[row];
        float* __restrict__ orow = plane + row * ld;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t t = t0 + lane + 64 * q + SW_S * r;
                const uint32_t h = acc16[k][q][r & 1];
                uint32_t v = (r < 2) ? (h & 0xffffu) : (h >> 16);
                if constexpr (USE32) v += acc32[k][q * 4 + r];
                if (t < ld) orow[t] = (t < len) ? static_cast<float>(v) : 0.f;
            }
    }
}

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

}  // namespace

struct spp_sweep_plan {
    int64_t ndm = 0, nchans = 0, nsamps = 0;
    int nbits = 8, big = 0;
    std::vector<int32_t> table;  // [ndm][nchans]
    std::vector<int32_t> tile_dm0, tile_ndm, tile_spread;
    std::vector<int32_t> base;   // [ntiles][nchans]
    std::vector<int32_t> row_len;
    int64_t max_len = 0, min_len = 0, total = 0;
    int32_t max_spread = 0;
    bool use32 = false;
    int64_t gpitch = 0;
    int64_t off_table = 0, off_base = 0, off_dm0 = 0, off_ndm = 0, off_spread = 0, off_rowlen = 0,
            off_g = 0, ws_bytes = 0;
    const void* uploaded_to = nullptr;
};

extern "C" int spp_sweep_plan_create(const int32_t* table_host, int64_t ndm, int64_t nchans,
                                     int64_t nsamps, int nbits, int big, spp_sweep_plan** plan) {
    spp::clear_error();
    SPP_REQUIRE(plan != nullptr, "null plan pointer");
    *plan = nullptr;
    SPP_REQUIRE(table_host != nullptr, "null delay table");
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8,
                "sweep needs nbits in {1, 2, 4, 8}, got %d", nbits);
    SPP_REQUIRE(ndm > 0 && nchans > 0 && nsamps > 0, "ndm, nchans, nsamps must be > 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(nsamps < (int64_t(1) << 31), "nsamps must be < 2^31");
    SPP_REQUIRE(ndm < (int64_t(1) << 31), "ndm must be < 2^31");
    const uint32_t maxv = (nbits == 8) ? 255u : ((1u << nbits) - 1u);
    SPP_REQUIRE(static_cast<double>(nchans) * maxv < 16777216.0,
                "nchans * max sample (%lld) must stay below 2^24 for exact float32 sums",
                static_cast<long long>(nchans * maxv));
    spp_sweep_plan* p = new (std::nothrow) spp_sweep_plan();
    SPP_REQUIRE(p != nullptr, "out of host memory");
    p->ndm = ndm;
    p->nchans = nchans;
    p->nsamps = nsamps;
    p->nbits = nbits;
    p->big = big ? 1 : 0;
    p->table.assign(table_host, table_host + ndm * nchans);
    p->row_len.resize(ndm);
    p->max_len = 0;
    p->min_len = nsamps;
    for (int64_t i = 0; i < ndm; ++i) {
        int32_t md = 0;
        const int32_t* r = &p->table[i * nchans];
        for (int64_t c = 0; c < nchans; ++c) {
            if (r[c] < 0) {
                delete p;
                SPP_REQUIRE(false, "sweep needs non-negative delays (row %lld channel %lld = %d)",
                            static_cast<long long>(i), static_cast<long long>(c), r[c]);
            }
            md = std::max(md, r[c]);
        }
        const int64_t len = nsamps - md;
        if (len <= 0) {
            delete p;
            SPP_REQUIRE(false, "DM row %lld: max delay %d >= nsamps %lld", static_cast<long long>(i),
                        md, static_cast<long long>(nsamps));
        }
        p->row_len[i] = static_cast<int32_t>(len);
        p->max_len = std::max(p->max_len, len);
        p->min_len = std::min(p->min_len, len);
        p->total += len;
    }
    // Greedy DM tiles: consecutive rows while the per-channel delay spread fits LDS.
    std::vector<int32_t> lo(nchans), hi(nchans);
    for (int64_t i = 0; i < ndm;) {
        const int32_t* r0 = &p->table[i * nchans];
        std::copy(r0, r0 + nchans, lo.begin());
        std::copy(r0, r0 + nchans, hi.begin());
        int32_t spread = 0;
        int64_t j = i + 1;
        while (j < ndm && j - i < SW_DT) {
            const int32_t* rj = &p->table[j * nchans];
            int32_t s = 0;
            for (int64_t c = 0; c < nchans; ++c)
                s = std::max(s, std::max(hi[c], rj[c]) - std::min(lo[c], rj[c]));
            if (s > SW_SPREAD_MAX) break;
            for (int64_t c = 0; c < nchans; ++c) {
                lo[c] = std::min(lo[c], rj[c]);
                hi[c] = std::max(hi[c], rj[c]);
            }
            spread = s;
            ++j;
        }
        p->tile_dm0.push_back(static_cast<int32_t>(i));
        p->tile_ndm.push_back(static_cast<int32_t>(j - i));
        p->tile_spread.push_back(spread);
        p->base.insert(p->base.end(), lo.begin(), lo.end());
        p->max_spread = std::max(p->max_spread, spread);
        i = j;
    }
    p->use32 = static_cast<int64_t>(nchans) * maxv > 65535;
    p->gpitch = ((nsamps + 63) / 64) * 64;
    const int64_t ntiles = static_cast<int64_t>(p->tile_dm0.size());
    int64_t off = 0;
    p->off_table = off; off = align256(off + 4 * ndm * nchans);
    p->off_base = off; off = align256(off + 4 * ntiles * nchans);
    p->off_dm0 = off; off = align256(off + 4 * ntiles);
    p->off_ndm = off; off = align256(off + 4 * ntiles);
    p->off_spread = off; off = align256(off + 4 * ntiles);
    p->off_rowlen = off; off = align256(off + 4 * ndm);
    p->off_g = off; off = align256(off + 4 * nchans * p->gpitch);
    p->ws_bytes = off;
    *plan = p;
    return SPP_OK;
}

extern "C" int spp_sweep_plan_destroy(spp_sweep_plan* plan) {
    delete plan;
    return SPP_OK;
}

extern "C" int spp_sweep_plan_info(const spp_sweep_plan* p, spp_sweep_info* info) {
    spp::clear_error();
    SPP_REQUIRE(p != nullptr && info != nullptr, "null argument");
    info->workspace_bytes = p->ws_bytes;
    info->max_len = p->max_len;
    info->min_len = p->min_len;
    info->total_samples = p->total;
    info->ndm_tiles = static_cast<int64_t>(p->tile_dm0.size());
    info->max_spread = p->max_spread;
    info->mode = (p->nbits == 8 ? 16 : 8) + (p->use32 ? 32 : 0);
    return SPP_OK;
}

extern "C" int spp_sweep_plan_row_lengths(const spp_sweep_plan* p, int64_t* out_host) {
    spp::clear_error();
    SPP_REQUIRE(p != nullptr && out_host != nullptr, "null argument");
    for (int64_t i = 0; i < p->ndm; ++i) out_host[i] = p->row_len[i];
    return SPP_OK;
}

extern "C" int spp_sweep_plan_upload(spp_sweep_plan* p, void* workspace, int64_t workspace_bytes,
                                     void* stream) {
    spp::clear_error();
    SPP_REQUIRE(p != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace_bytes >= p->ws_bytes, "workspace too small: %lld < %lld bytes",
                static_cast<long long>(workspace_bytes), static_cast<long long>(p->ws_bytes));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    const int64_t ntiles = static_cast<int64_t>(p->tile_dm0.size());
    SPP_HIP(hipMemcpyAsync(ws + p->off_table, p->table.data(), 4 * p->ndm * p->nchans,
                           hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + p->off_base, p->base.data(), 4 * ntiles * p->nchans,
                           hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + p->off_dm0, p->tile_dm0.data(), 4 * ntiles, hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + p->off_ndm, p->tile_ndm.data(), 4 * ntiles, hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + p->off_spread, p->tile_spread.data(), 4 * ntiles,
                           hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + p->off_rowlen, p->row_len.data(), 4 * p->ndm, hipMemcpyHostToDevice, st));
    SPP_HIP(hipStreamSynchronize(st));
    p->uploaded_to = workspace;
    return SPP_OK;
}

namespace {

template <int NBITS, bool BIG>
int launch_build(const uint8_t* packed, const spp_sweep_plan* p, uint32_t* G, hipStream_t st) {
    dim3 grid(static_cast<unsigned>(spp::ceil_div(p->gpitch, G_BUILD_T)),
              static_cast<unsigned>(spp::ceil_div(p->nchans, G_BUILD_C)));
    hipLaunchKernelGGL((k_build_g<NBITS, BIG>), grid, dim3(256), 0, st, packed, p->nsamps, p->nchans, G,
                       p->gpitch);
    SPP_LAUNCHED("k_build_g");
    return SPP_OK;
}

template <int NBITS, bool USE32>
int launch_sweep(const spp_sweep_plan* p, char* ws, float* plane, int64_t ld, hipStream_t st) {
    const int64_t ntiles_dm = static_cast<int64_t>(p->tile_dm0.size());
    const int64_t ntiles_t = spp::ceil_div(p->max_len, SW_T);
    const int64_t nblocks = ntiles_dm * ntiles_t;
    const int64_t grid = spp::ceil_div(nblocks, 8) * 8;
    SPP_REQUIRE(grid <= 0x7fffffffLL, "sweep grid too large");
    hipLaunchKernelGGL((k_sweep<NBITS, USE32>), dim3(static_cast<unsigned>(grid)), dim3(SW_THREADS), 0,
                       st, reinterpret_cast<const uint32_t*>(ws + p->off_g), p->gpitch, p->nchans,
                       reinterpret_cast<const int32_t*>(ws + p->off_table),
                       reinterpret_cast<const int32_t*>(ws + p->off_base),
                       reinterpret_cast<const int32_t*>(ws + p->off_dm0),
                       reinterpret_cast<const int32_t*>(ws + p->off_ndm),
                       reinterpret_cast<const int32_t*>(ws + p->off_spread),
                       reinterpret_cast<const int32_t*>(ws + p->off_rowlen), plane, ld, ntiles_dm,
                       nblocks);
    SPP_LAUNCHED("k_sweep");
    return SPP_OK;
}

}  // namespace

namespace {

int check_exec(const spp_sweep_plan* p, const void* workspace, int64_t workspace_bytes) {
    SPP_REQUIRE(p != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace == p->uploaded_to, "workspace is not the buffer the plan was uploaded to");
    SPP_REQUIRE(workspace_bytes >= p->ws_bytes, "workspace too small");
    return SPP_OK;
}

}  // namespace

extern "C" int spp_sweep_build(const spp_sweep_plan* p, const uint8_t* packed, void* workspace,
                               int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    SPP_REQUIRE(packed != nullptr, "null packed stream");
    hipStream_t st = spp::as_stream(stream);
    uint32_t* G = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + p->off_g);
    switch (p->nbits * 2 + p->big) {
        case 2: return launch_build<1, false>(packed, p, G, st);
        case 3: return launch_build<1, true>(packed, p, G, st);
        case 4: return launch_build<2, false>(packed, p, G, st);
        case 5: return launch_build<2, true>(packed, p, G, st);
        case 8: return launch_build<4, false>(packed, p, G, st);
        case 9: return launch_build<4, true>(packed, p, G, st);
        default: return launch_build<8, false>(packed, p, G, st);
    }
}

extern "C" int spp_sweep_run(const spp_sweep_plan* p, float* plane, int64_t ld, void* workspace,
                             int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    SPP_REQUIRE(plane != nullptr, "null plane");
    SPP_REQUIRE(ld >= p->max_len, "ld (%lld) < longest row (%lld)", static_cast<long long>(ld),
                static_cast<long long>(p->max_len));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    switch (p->nbits) {
        case 1: return p->use32 ? launch_sweep<1, true>(p, ws, plane, ld, st) : launch_sweep<1, false>(p, ws, plane, ld, st);
        case 2: return p->use32 ? launch_sweep<2, true>(p, ws, plane, ld, st) : launch_sweep<2, false>(p, ws, plane, ld, st);
        case 4: return p->use32 ? launch_sweep<4, true>(p, ws, plane, ld, st) : launch_sweep<4, false>(p, ws, plane, ld, st);
        default: return p->use32 ? launch_sweep<8, true>(p, ws, plane, ld, st) : launch_sweep<8, false>(p, ws, plane, ld, st);
    }
}

extern "C" int spp_sweep_execute(const spp_sweep_plan* p, const uint8_t* packed, float* plane,
                                 int64_t ld, void* workspace, int64_t workspace_bytes,
                                 void* stream) {
    int rc = spp_sweep_build(p, packed, workspace, workspace_bytes, stream);
    if (rc) return rc;
    return spp_sweep_run(p, plane, ld, workspace, workspace_bytes, stream);
}
