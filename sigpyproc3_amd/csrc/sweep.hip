// DM-trial sweep of one plan: plane[i][t] = sum_c x_c[t + d_i(c)] for every DM row i of a
// delay table, over a whole stream resident in HBM. Stage 1 (k_build_g) writes the
// channel-major interleaved copy G of the packed stream; stage 2 (k_sweep) sums DM tiles from
// it. The device code of both stages is in spp_sweep_dev.hpp (shared with the batch kernel);
// this file holds the one-plan kernels, the host planner and the C ABI.
#include "spp_sweep_dev.hpp"
#include "spp_sweep_plan.hpp"

#include <algorithm>
#include <new>
#include <vector>

namespace {

constexpr int G_BCW = 64;  // packed bytes per row a stage-1 block takes

// Stage 1. Block (bx, by): words [bx*B_TJ, bx*B_TJ + B_TJ) of the channels in row bytes
// [by*64, by*64 + 64) (64 * 8/NBITS channels). Blocks are numbered with by fastest and
// consecutive numbers on one XCD, so the column blocks of a row range (which share its
// 128-byte lines) and the blocks bx, bx + 2, ..., bx + 2(IL - 1) (which stage the same rows,
// S being 2 * B_TJ) run close together and re-read those rows from that XCD's L2.
template <int NBITS, bool BIG, bool ALIGNED>
__global__ __launch_bounds__(B_THREADS) void k_build_g(const uint8_t* __restrict__ in, int64_t nsamps,
                                                       int64_t nchans, uint32_t* __restrict__ G,
                                                       int64_t gpitch, int64_t glen) {
    static_assert(SW_S == 2 * B_TJ, "stage-1 blocks two apart share staged rows");
    __shared__ uint8_t raw[build_lds_bytes<NBITS, G_BCW>()];
    const int64_t row_bytes = nchans * NBITS / 8;
    const int64_t nby = (row_bytes + G_BCW - 1) / G_BCW;
    const int64_t b = blockIdx.x;
    const int64_t logical = (b % 8) * (gridDim.x / 8) + b / 8;
    if (logical >= nby * ((glen + B_TJ - 1) / B_TJ)) return;
    build_tile<NBITS, BIG, ALIGNED, G_BCW>(in, nsamps, nchans, G, gpitch, glen, (logical / nby) * B_TJ,
                                           (logical % nby) * G_BCW, raw, threadIdx.x);
}

// Stage 2: one block per (time tile, DM tile, sub-tile of K * WAVES rows). XCD-aware order:
// consecutive blocks on one XCD walk the DM tiles of one time tile.
template <int NBITS, bool USE32, int K, int WAVES, int CC, int MINW>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(MINW))) void k_sweep(
    const uint32_t* __restrict__ G, int64_t gpitch, int64_t glen, int64_t nchans,
    const uint32_t* __restrict__ relw4, int64_t rpitch4, const int32_t* __restrict__ basew,
    const int32_t* __restrict__ tile_dm0, const int32_t* __restrict__ tile_ndm,
    const int32_t* __restrict__ tile_spread, const int32_t* __restrict__ row_len,
    float* __restrict__ plane, int64_t ld, int64_t shrink, int64_t ntiles_dm, int64_t nblocks) {
    constexpr int NSUB = SW_DT / (K * WAVES);
    constexpr int T = Fields<NBITS>::IL * SW_S;
    __shared__ uint32_t lds[2 * CC * SW_LW];
    const int64_t b = blockIdx.x;
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t logical = (b % 8) * per_xcd + b / 8;
    if (logical >= nblocks) return;
    const int64_t u = logical % (ntiles_dm * NSUB);
    const int64_t dt = u / NSUB;
    const int sub = static_cast<int>(u % NSUB);
    const int64_t t0 = (logical / (ntiles_dm * NSUB)) * T;
#define SPP_TILE_EXIT return
#define SPP_TILE_TID threadIdx.x
#include "spp_sweep_tile.inc"
#undef SPP_TILE_TID
#undef SPP_TILE_EXIT
}

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

}  // namespace


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
    SPP_REQUIRE(nsamps < (int64_t(1) << 30), "nsamps must be < 2^30");
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
    p->rpitch = ((nchans + 64 + 15) / 16) * 16;
    p->relw.assign(static_cast<size_t>(ndm * p->rpitch), 0);
    static_assert(SW_SPREAD_MAX + 3 < 256, "window offsets fit one byte");
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
        for (int64_t r = i; r < j; ++r)
            for (int64_t c = 0; c < nchans; ++c)
                p->relw[r * p->rpitch + c] =
                    static_cast<uint8_t>(p->table[r * nchans + c] - (lo[c] & ~int32_t(3)));
        for (int64_t c = 0; c < nchans; ++c) p->base.push_back(lo[c] & ~int32_t(3));
        p->max_spread = std::max(p->max_spread, spread);
        i = j;
    }
    p->use32 = static_cast<int64_t>(nchans) * maxv > 65535;
    // Nibble data: 16 waves x 1 DM row x 16 channels per stage at 8 waves per SIMD (variant 4);
    // byte data: 8 waves x 2 DM rows x 12 channels at 6 waves per SIMD (variant 0); tunable
    p->variant = nbits <= 2 ? 4 : 0;
    p->gpitch = ((nsamps + 63) / 64) * 64;
    const int64_t ntiles = static_cast<int64_t>(p->tile_dm0.size());
    int64_t off = 0;
    p->off_relw = off; off = align256(off + ndm * p->rpitch);
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

extern "C" int spp_sweep_variant_count(void) { return 6; }

extern "C" int spp_sweep_plan_set_variant(spp_sweep_plan* p, int variant) {
    spp::clear_error();
    SPP_REQUIRE(p != nullptr, "null plan");
    SPP_REQUIRE(variant >= 0 && variant < spp_sweep_variant_count(), "variant %d outside [0, %d)", variant,
                spp_sweep_variant_count());
    p->variant = variant;
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
    info->mode = (p->nbits <= 2 ? 4 : (p->nbits == 4 ? 8 : 16)) + (p->use32 ? 32 : 0) + 64 * p->variant;
    info->tables_bytes = p->off_g;
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
    SPP_REQUIRE(workspace_bytes >= p->off_g, "workspace too small: %lld < %lld bytes (tables)",
                static_cast<long long>(workspace_bytes), static_cast<long long>(p->off_g));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    const int64_t ntiles = static_cast<int64_t>(p->tile_dm0.size());
    SPP_HIP(hipMemcpyAsync(ws + p->off_relw, p->relw.data(), p->ndm * p->rpitch,
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
    p->uploaded_with_g = workspace_bytes >= p->ws_bytes;
    return SPP_OK;
}

namespace {

// One execution of a plan on a stream of n <= plan nsamps samples: every row is shorter by
// shrink = nsamps - n; the stage-1 copy G (gpitch words per channel) holds glen valid words
// per channel.
struct Extent {
    int64_t n, glen, shrink, max_len;
    uint32_t* G;
    int64_t gpitch;
};

// G inside the plan's workspace (pitch for the plan's nsamps).
Extent extent_of(const spp_sweep_plan* p, int64_t n, void* workspace) {
    const int64_t glen = std::min(p->gpitch, ((n + 63) / 64) * 64);
    return Extent{n, glen, p->nsamps - n, p->max_len - (p->nsamps - n),
                  reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + p->off_g), p->gpitch};
}

// G in a caller's buffer, pitched for n samples.
Extent extent_own_g(const spp_sweep_plan* p, int64_t n, uint32_t* G) {
    const int64_t glen = ((n + 63) / 64) * 64;
    return Extent{n, glen, p->nsamps - n, p->max_len - (p->nsamps - n), G, glen};
}

template <int NBITS, bool BIG>
int launch_build(const uint8_t* packed, const spp_sweep_plan* p, const Extent& e, hipStream_t st) {
    const int64_t row_bytes = p->nchans * NBITS / 8;
    const int64_t nblocks = spp::ceil_div(e.glen, int64_t(B_TJ)) * spp::ceil_div(row_bytes, int64_t(G_BCW));
    SPP_REQUIRE(nblocks + 8 <= 0x7fffffffLL, "stage-1 grid too large");
    const dim3 grid(static_cast<unsigned>(spp::ceil_div(nblocks, int64_t(8)) * 8));
    if (row_bytes % 4 == 0 && reinterpret_cast<uintptr_t>(packed) % 4 == 0)
        hipLaunchKernelGGL((k_build_g<NBITS, BIG, true>), grid, dim3(B_THREADS), 0, st, packed, e.n, p->nchans, e.G,
                           e.gpitch, e.glen);
    else
        hipLaunchKernelGGL((k_build_g<NBITS, BIG, false>), grid, dim3(B_THREADS), 0, st, packed, e.n, p->nchans,
                           e.G, e.gpitch, e.glen);
    SPP_LAUNCHED("k_build_g");
    return SPP_OK;
}

template <int NBITS, bool USE32, int K, int WAVES, int CC, int MINW>
int launch_sweep_v(const spp_sweep_plan* p, const Extent& e, char* ws, float* plane, int64_t ld,
                   hipStream_t st) {
    const int64_t ntiles_dm = static_cast<int64_t>(p->tile_dm0.size());
    const int64_t ntiles_t = spp::ceil_div(e.max_len, int64_t(Fields<NBITS>::IL) * SW_S);
    const int64_t nblocks = ntiles_dm * (SW_DT / (K * WAVES)) * ntiles_t;
    const int64_t grid = spp::ceil_div(nblocks, 8) * 8;
    SPP_REQUIRE(grid <= 0x7fffffffLL, "sweep grid too large");
    hipLaunchKernelGGL((k_sweep<NBITS, USE32, K, WAVES, CC, MINW>), dim3(static_cast<unsigned>(grid)),
                       dim3(WAVES * 64), 0, st, e.G, e.gpitch, e.glen, p->nchans, reinterpret_cast<const uint32_t*>(ws + p->off_relw),
                       p->rpitch / 4, reinterpret_cast<const int32_t*>(ws + p->off_base),
                       reinterpret_cast<const int32_t*>(ws + p->off_dm0),
                       reinterpret_cast<const int32_t*>(ws + p->off_ndm),
                       reinterpret_cast<const int32_t*>(ws + p->off_spread),
                       reinterpret_cast<const int32_t*>(ws + p->off_rowlen), plane, ld, e.shrink,
                       ntiles_dm, nblocks);
    SPP_LAUNCHED("k_sweep");
    return SPP_OK;
}

// Block shapes: (DM rows per wave, waves, channels per stage, min waves per SIMD); all tile
// SW_DT rows and take a multiple of 4 channels per stage (one packed offset word per 4).
// The kernel is bound by its LDS reads. One DM row per wave in 16-wave blocks holds the
// accumulators of one row (about 60 VGPRs, no spills) and so runs 8 waves per SIMD; with 16
// channels per stage (64 KB of LDS, two blocks per CU) it is the fastest shape measured for
// nibble data and on a par with the 2 x 8 x 12 shape for byte data.
template <int NBITS, bool USE32>
int launch_sweep(const spp_sweep_plan* p, const Extent& e, char* ws, float* plane, int64_t ld,
                 hipStream_t st) {
    if constexpr (NBITS <= 2) {
        switch (p->variant) {
            case 1: return launch_sweep_v<NBITS, USE32, 2, 8, 20, 4>(p, e, ws, plane, ld, st);
            case 2: return launch_sweep_v<NBITS, USE32, 2, 8, 8, 6>(p, e, ws, plane, ld, st);
            case 3: return launch_sweep_v<NBITS, USE32, 2, 8, 12, 6>(p, e, ws, plane, ld, st);
            case 4: return launch_sweep_v<NBITS, USE32, 1, 16, 16, 8>(p, e, ws, plane, ld, st);
            case 5: return launch_sweep_v<NBITS, USE32, 1, 16, 12, 8>(p, e, ws, plane, ld, st);
            default: return launch_sweep_v<NBITS, USE32, 2, 8, 16, 4>(p, e, ws, plane, ld, st);
        }
    } else {
        switch (p->variant) {
            case 1: return launch_sweep_v<NBITS, USE32, 2, 8, 16, 4>(p, e, ws, plane, ld, st);
            case 2: return launch_sweep_v<NBITS, USE32, 4, 4, 8, 3>(p, e, ws, plane, ld, st);
            case 3: return launch_sweep_v<NBITS, USE32, 2, 8, 8, 4>(p, e, ws, plane, ld, st);
            case 4: return launch_sweep_v<NBITS, USE32, 1, 16, 16, 8>(p, e, ws, plane, ld, st);
            case 5: return launch_sweep_v<NBITS, USE32, 1, 16, 12, 8>(p, e, ws, plane, ld, st);
            default: return launch_sweep_v<NBITS, USE32, 2, 8, 12, 6>(p, e, ws, plane, ld, st);
        }
    }
}

}  // namespace

namespace {

int check_exec(const spp_sweep_plan* p, const void* workspace, int64_t workspace_bytes, bool own_g = false) {
    SPP_REQUIRE(p != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace == p->uploaded_to, "workspace is not the buffer the plan was uploaded to");
    SPP_REQUIRE(workspace_bytes >= (own_g ? p->off_g : p->ws_bytes), "workspace too small");
    SPP_REQUIRE(own_g || p->uploaded_with_g, "the plan was uploaded without room for its stage-1 copy");
    return SPP_OK;
}

int build_n(const spp_sweep_plan* p, const uint8_t* packed, const Extent& e, hipStream_t st) {
    switch (p->nbits * 2 + p->big) {
        case 2: return launch_build<1, false>(packed, p, e, st);
        case 3: return launch_build<1, true>(packed, p, e, st);
        case 4: return launch_build<2, false>(packed, p, e, st);
        case 5: return launch_build<2, true>(packed, p, e, st);
        case 8: return launch_build<4, false>(packed, p, e, st);
        case 9: return launch_build<4, true>(packed, p, e, st);
        default: return launch_build<8, false>(packed, p, e, st);
    }
}

int run_n(const spp_sweep_plan* p, const Extent& e, float* plane, int64_t ld, void* workspace,
          hipStream_t st) {
    SPP_REQUIRE(plane != nullptr, "null plane");
    SPP_REQUIRE(ld >= e.max_len, "ld (%lld) < longest row (%lld)", static_cast<long long>(ld),
                static_cast<long long>(e.max_len));
    char* ws = static_cast<char*>(workspace);
    switch (p->nbits) {
        case 1: return p->use32 ? launch_sweep<1, true>(p, e, ws, plane, ld, st) : launch_sweep<1, false>(p, e, ws, plane, ld, st);
        case 2: return p->use32 ? launch_sweep<2, true>(p, e, ws, plane, ld, st) : launch_sweep<2, false>(p, e, ws, plane, ld, st);
        case 4: return p->use32 ? launch_sweep<4, true>(p, e, ws, plane, ld, st) : launch_sweep<4, false>(p, e, ws, plane, ld, st);
        default: return p->use32 ? launch_sweep<8, true>(p, e, ws, plane, ld, st) : launch_sweep<8, false>(p, e, ws, plane, ld, st);
    }
}

}  // namespace

extern "C" int spp_sweep_build(const spp_sweep_plan* p, const uint8_t* packed, void* workspace,
                               int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    SPP_REQUIRE(packed != nullptr, "null packed stream");
    return build_n(p, packed, extent_of(p, p->nsamps, workspace), spp::as_stream(stream));
}

extern "C" int spp_sweep_run(const spp_sweep_plan* p, float* plane, int64_t ld, void* workspace,
                             int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    return run_n(p, extent_of(p, p->nsamps, workspace), plane, ld, workspace, spp::as_stream(stream));
}

extern "C" int spp_sweep_execute(const spp_sweep_plan* p, const uint8_t* packed, float* plane,
                                 int64_t ld, void* workspace, int64_t workspace_bytes,
                                 void* stream) {
    int rc = spp_sweep_build(p, packed, workspace, workspace_bytes, stream);
    if (rc) return rc;
    return spp_sweep_run(p, plane, ld, workspace, workspace_bytes, stream);
}

extern "C" int spp_sweep_execute_n(const spp_sweep_plan* p, const uint8_t* packed, int64_t nsamps,
                                   float* plane, int64_t ld, void* workspace,
                                   int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    SPP_REQUIRE(packed != nullptr, "null packed stream");
    SPP_REQUIRE(nsamps <= p->nsamps, "nsamps %lld exceeds the plan's %lld", static_cast<long long>(nsamps),
                static_cast<long long>(p->nsamps));
    SPP_REQUIRE(nsamps > p->nsamps - p->min_len,
                "nsamps %lld must exceed the largest delay of the table (%lld)",
                static_cast<long long>(nsamps), static_cast<long long>(p->nsamps - p->min_len));
    const Extent e = extent_of(p, nsamps, workspace);
    hipStream_t st = spp::as_stream(stream);
    rc = build_n(p, packed, e, st);
    if (rc) return rc;
    return run_n(p, e, plane, ld, workspace, st);
}

extern "C" int spp_sweep_g_bytes(const spp_sweep_plan* p, int64_t nsamps, int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(p != nullptr && bytes != nullptr, "null argument");
    SPP_REQUIRE(nsamps > 0 && nsamps <= p->nsamps, "nsamps %lld outside (0, %lld]", static_cast<long long>(nsamps),
                static_cast<long long>(p->nsamps));
    *bytes = 4 * p->nchans * (((nsamps + 63) / 64) * 64);
    return SPP_OK;
}

extern "C" int spp_sweep_execute_g(const spp_sweep_plan* p, const uint8_t* packed, int64_t nsamps, float* plane,
                                   int64_t ld, void* tables, int64_t tables_bytes, uint32_t* g, int64_t g_bytes,
                                   void* stream) {
    spp::clear_error();
    int rc = check_exec(p, tables, tables_bytes, true);
    if (rc) return rc;
    SPP_REQUIRE(packed != nullptr && g != nullptr, "null argument");
    SPP_REQUIRE(nsamps <= p->nsamps, "nsamps %lld exceeds the plan's %lld", static_cast<long long>(nsamps),
                static_cast<long long>(p->nsamps));
    SPP_REQUIRE(nsamps > p->nsamps - p->min_len,
                "nsamps %lld must exceed the largest delay of the table (%lld)",
                static_cast<long long>(nsamps), static_cast<long long>(p->nsamps - p->min_len));
    const Extent e = extent_own_g(p, nsamps, g);
    SPP_REQUIRE(g_bytes >= 4 * p->nchans * e.gpitch, "stage-1 buffer too small: %lld < %lld bytes",
                static_cast<long long>(g_bytes), static_cast<long long>(4 * p->nchans * e.gpitch));
    hipStream_t st = spp::as_stream(stream);
    rc = build_n(p, packed, e, st);
    if (rc) return rc;
    return run_n(p, e, plane, ld, tables, st);
}
