// One-DM dedispersion and subbanding of a gulp, with the sub-byte unpack fused into
// the LDS staging (the "one DM per launch" form).
//
// Gulp layout: file order, [nsamps][nchans] packed nbits (1/2/4/8, bit order `big`)
// or float32. A block owns DD_T consecutive output samples and a group of channels.
// For each stage it picks the longest run of channels (at most 16 or DD_CC, the tile
// width) whose delay spread fits the LDS rows, stages rows [t0 + dmin, t0 + DD_T + dmax)
// of those channels as one byte per sample (rows of an odd number of dwords, so the
// per-sample column reads below are bank-conflict free), and every thread adds its two
// outputs' samples. Integer sums are exact (every partial sum < 2^24), so splitting channels
// across blocks and combining with float atomics is bit-exact and order-free.
//
// Work per output sample: nchans adds; HBM bytes per gulp: nsamps * nchans * nbits/8
// read + 4 * nout written. For 1/2-bit data that is under a byte per add, and the kernel is
// bound by instruction issue (a byte write per sample staged, a byte read and an add per
// sample summed), not by HBM.
#include "spp_common.hpp"
#include "spp_gulp_swar.hpp"
#include "spp_unpack.hpp"

#include <algorithm>
#include <type_traits>

namespace {

constexpr int DD_THREADS = 256;
constexpr int DD_T = 512;       // output samples per block (2 per thread)
constexpr int DD_CC = 64;       // channels per stage at most
constexpr int DD_ROWS = 768;    // staged rows per stage: spread <= DD_ROWS - DD_T
constexpr int DD_ROWS_LOW = 576;  // the same for small delay spreads (four 64-channel tiles per CU)
constexpr int DD_LD = 16;       // staging loads in flight per thread
constexpr int DDF_CC = 16;      // float32 path: channels per stage
constexpr int DDF_ROWS = 768;
constexpr int DDF_STRIDE = 17;  // floats per staged row

// Greedy stage: the longest prefix of sdel[0..ncand) (ncand <= 64) whose delay spread is at
// most max_spread, with its minimum and maximum delay. Every wave computes it: an inclusive
// scan of min and max across the lanes, and since the prefix spreads only grow, the lanes
// whose prefix fits are lanes 0 .. n-1.
__device__ __forceinline__ int stage_extent(const int32_t* sdel, int ncand, int max_spread,
                                            int& dmin, int& dmax) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int d = sdel[lane < ncand ? lane : 0];
    int mn = d, mx = d;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int a = __shfl_up(mn, off), b = __shfl_up(mx, off);
        if (lane >= off) {
            mn = min(mn, a);
            mx = max(mx, b);
        }
    }
    const bool ok = lane < ncand && mx - mn <= max_spread;
    const int n = __popcll(__ballot(ok));
    dmin = __shfl(mn, n - 1);
    dmax = __shfl(mx, n - 1);
    return n;
}

// MODE 0: dedisperse (out[out_start + t] = sum). MODE 1: subband (out[t*nsubs + sub] += run sum).
// Staging: the stage's channels occupy bytes [bb0, bb1) of each packed row. All threads take
// (row, unit) elements, a unit being a dword (WIDE: rows and the stream start are 4-byte
// aligned) or a byte, DD_LD of them in flight per thread, and unpack every sample of a unit
// that belongs to the stage into its tile column.
template <int NBITS, bool BIG, int MODE, bool WIDE, int CCX, int ROWS>
__global__ __launch_bounds__(DD_THREADS) void k_gulp_int(
    const uint8_t* __restrict__ in, const int32_t* __restrict__ delays,
    const int32_t* __restrict__ chan_to_sub, float* __restrict__ out, int64_t nchans, int64_t nsamps,
    int64_t nout, int64_t out_start, int64_t row_bytes, int64_t chans_per_group, int use_atomic,
    int64_t nsubs, int64_t out_size, int64_t ngroups, int64_t nblocks, int64_t ilv, int64_t nchunk) {
    // CCX channels per stage at most; rows of CCX + 4 bytes (an odd number of dwords, so
    // the per-sample column reads are bank-conflict free). Narrow channel groups use the
    // small tile and fit more blocks per CU.
    constexpr int STRIDE = CCX + 4;
    static_assert((STRIDE / 4) % 2 == 1 && STRIDE % 4 == 0, "odd dword row stride");
    __shared__ __align__(16) uint8_t tile[ROWS * STRIDE];
    __shared__ __align__(16) int32_t sdel[CCX];
    __shared__ __align__(16) int32_t ssub[CCX];
    __shared__ __align__(16) int32_t soff[CCX];  // tile offset of channel k's column at row rel
    __shared__ int32_t sfast[CCX];  // MODE 1: band of channels k .. k+7 if they share one, else -2
    // XCD-aware order (blocks go to the 8 XCDs round-robin): consecutive logical blocks run on
    // one XCD and are the channel groups of one time tile, so the groups that share a packed
    // row's cache lines read them from that XCD's L2 at about the same time.
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (logical >= nblocks) return;
    const int tid = threadIdx.x;
    const int64_t t0 = (logical / ngroups) * DD_T;
    // The block's channels: nchunk chunks of chans_per_group, chunk m starting at
    // ((gh * nchunk + m) * ilv + gl) * chans_per_group for group g = gh * ilv + gl. With ilv > 1
    // the ilv groups of a time tile that read one packed row line run side by side.
    const int64_t grp = logical % ngroups;
    const int64_t gl = grp % ilv, gh = grp / ilv;
    const int64_t ta = t0 + tid;
    const int64_t tb = t0 + DD_THREADS + tid;
    uint32_t acc0 = 0, acc1 = 0;
    int cur_sub = -1;

    auto flush_sub = [&](int sub) {
        if (sub >= 0) {
            const int64_t ia = ta * nsubs + sub, ib = tb * nsubs + sub;
            if (ta < nout && ia < out_size) atomicAdd(out + ia, static_cast<float>(acc0));
            if (tb < nout && ib < out_size) atomicAdd(out + ib, static_cast<float>(acc1));
        }
        acc0 = 0;
        acc1 = 0;
    };

    for (int64_t m = 0; m < nchunk; ++m) {
        const int64_t cb = ((gh * nchunk + m) * ilv + gl) * chans_per_group;
        if (cb >= nchans) break;
        const int64_t ce = min(nchans, cb + chans_per_group);
        for (int64_t c = cb; c < ce;) {
            const int ncand = static_cast<int>(min<int64_t>(CCX, ce - c));
            __syncthreads();
            if (tid < ncand) {
                sdel[tid] = delays[c + tid];
                if constexpr (MODE == 1) ssub[tid] = chan_to_sub[c + tid];
            }
            __syncthreads();
            int dmin, dmax;
            int n = stage_extent(sdel, ncand, ROWS - DD_T, dmin, dmax);
            if (n < ncand && n >= 4) {
                // Keep stage starts on multiples of 4 channels (dword-aligned tile columns for the
                // word unpack below); a prefix of a stage fits the rows too. dmin/dmax stay bounds.
                n &= ~3;
            }
            const int nrows = DD_T + (dmax - dmin);
            const int64_t rbase = t0 + dmin;
            if (tid < n) {  // read after the staging barrier
                soff[tid] = (sdel[tid] - dmin) * STRIDE + tid;
                if constexpr (MODE == 1) {
                    int f = -2;
                    if (tid + 8 <= n) {
                        f = ssub[tid];
    #pragma unroll
                        for (int j = 1; j < 8; ++j) f = ssub[tid + j] == f ? f : -2;
                    }
                    sfast[tid] = f;
                }
            }
            {
                constexpr int PER = 8 / NBITS;  // samples per byte
                constexpr int UB = WIDE ? 4 : 1;  // bytes per unit
                const int64_t bb0 = c * NBITS / 8, bb1 = ((c + n) * NBITS + 7) / 8;
                const int64_t ub0 = bb0 / UB * UB;  // first unit's byte
                const int nunit = static_cast<int>((bb1 - ub0 + UB - 1) / UB);
                const int col_base = static_cast<int>(ub0 * PER - c);  // tile column of unit 0's first sample
                // One staged unit (row r of the stage, unit q): interior dwords expand to 4 * PER
                // tile bytes with dword stores; edge units go byte by byte.
                auto put = [&](uint32_t v, int r, int q) {
                    const int col0 = col_base + UB * PER * q;  // tile column of the unit's first sample
                    if constexpr (WIDE) {
                        if (col0 >= 0 && col0 + 4 * PER <= n && (col0 & 3) == 0) {
                            uint32_t o[PER];
                            spp::expand_samples<NBITS, BIG>(v, o);
                            uint32_t* dst = reinterpret_cast<uint32_t*>(tile + r * STRIDE + col0);
    #pragma unroll
                            for (int j = 0; j < PER; ++j) dst[j] = o[j];
                            return;
                        }
                    }
    #pragma unroll
                    for (int jb = 0; jb < UB; ++jb) {
                        const uint32_t byte = (v >> (8 * jb)) & 0xffu;
    #pragma unroll
                        for (int k = 0; k < PER; ++k) {
                            const int ch = col0 + jb * PER + k;
                            if (ch >= 0 && ch < n) {
                                const int shift = (NBITS == 8) ? 0 : (BIG ? (8 - (k + 1) * NBITS) : k * NBITS);
                                tile[r * STRIDE + ch] = static_cast<uint8_t>((byte >> shift) & ((1u << NBITS) - 1u));
                            }
                        }
                    }
                };
                auto load = [&](const uint8_t* src) -> uint32_t {
                    return WIDE ? *reinterpret_cast<const uint32_t*>(src) : static_cast<uint32_t>(*src);
                };
                // Rows inside the gulp (no wrap, 32-bit offsets) with a power-of-two unit count: each
                // thread keeps one unit column and walks rows tid / nunit, + 256 / nunit, ...
                const bool inrange = rbase >= 0 && rbase + nrows <= nsamps &&
                                     static_cast<int64_t>(nrows) * row_bytes < (int64_t(1) << 31);
                if (inrange && (nunit & (nunit - 1)) == 0) {
                    const int lg = __builtin_ctz(static_cast<unsigned>(nunit));
                    const int q = tid & (nunit - 1);
                    const int rstep = DD_THREADS >> lg;
                    const int rb32 = static_cast<int>(row_bytes);
                    const uint8_t* sbase = in + rbase * row_bytes + ub0 + UB * q;
                    for (int r0 = tid >> lg; r0 < nrows; r0 += DD_LD * rstep) {
                        uint32_t v[DD_LD];
    #pragma unroll
                        for (int u = 0; u < DD_LD; ++u) {
                            const int r = r0 + u * rstep;
                            v[u] = r < nrows ? load(sbase + r * rb32) : 0u;
                        }
    #pragma unroll
                        for (int u = 0; u < DD_LD; ++u) {
                            const int r = r0 + u * rstep;
                            if (r < nrows) put(v[u], r, q);
                        }
                    }
                } else {
                    // Any unit count: element e is (row e / nunit, unit e % nunit), the division as a
                    // multiply-high (exact while e * nunit < 2^32: e < 768 * 65 here; magic 0 and the
                    // added e give e / 1).
                    const int nel = nrows * nunit;
                    const uint32_t magic = nunit > 1 ? static_cast<uint32_t>(0xffffffffu / static_cast<uint32_t>(nunit) + 1u) : 0u;
                    const uint32_t unit1 = nunit == 1 ? 1u : 0u;
                    for (int e0 = 0; e0 < nel; e0 += DD_LD * DD_THREADS) {
                        uint32_t v[DD_LD];
    #pragma unroll
                        for (int u = 0; u < DD_LD; ++u) {
                            const int e = e0 + u * DD_THREADS + tid;
                            const int r = static_cast<int>(__umulhi(static_cast<uint32_t>(e), magic) + unit1 * static_cast<uint32_t>(e));
                            const int q = e - r * nunit;
                            int64_t row = rbase + r;
                            if (row < 0) row += nsamps;  // numba wraparound of a negative index
                            v[u] = (e < nel && row >= 0 && row < nsamps) ? load(in + row * row_bytes + ub0 + UB * q) : 0u;
                        }
    #pragma unroll
                        for (int u = 0; u < DD_LD; ++u) {
                            const int e = e0 + u * DD_THREADS + tid;
                            const int r = static_cast<int>(__umulhi(static_cast<uint32_t>(e), magic) + unit1 * static_cast<uint32_t>(e));
                            if (e < nel) put(v[u], r, e - r * nunit);
                        }
                    }
                }
            }
            __syncthreads();
            // Channels in steps of 8 (their tile offsets read as two 16-byte broadcasts, 16 tile reads in
            // flight) while the 8 share the current subband; single channels otherwise. The step
            // choice depends only on LDS values every thread sees, so it is uniform.
            const int base0 = tid * STRIDE;  // + soff[k]: row tid + rel, column k
            const int base1 = base0 + DD_THREADS * STRIDE;
            int k = 0;
            while (k < n) {
                bool step8 = k + 8 <= n;
                if constexpr (MODE == 1) step8 = sfast[k] == cur_sub;  // -2 never equals a band
                if (step8) {
                    const int4 d0 = *reinterpret_cast<const int4*>(soff + k);
                    const int4 d1 = *reinterpret_cast<const int4*>(soff + k + 4);
                    const int dd[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
                    uint32_t a[8], b[8];
    #pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        a[j] = tile[base0 + dd[j]];
                        b[j] = tile[base1 + dd[j]];
                    }
    #pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        acc0 += a[j];
                        acc1 += b[j];
                    }
                    k += 8;
                } else {
                    if constexpr (MODE == 1) {
                        const int sub = ssub[k];
                        if (sub != cur_sub) {
                            flush_sub(cur_sub);
                            cur_sub = sub;
                            continue;  // channel k again, now possibly as the start of a step of 8
                        }
                    }
                    const int d = soff[k];
                    acc0 += tile[base0 + d];
                    acc1 += tile[base1 + d];
                    ++k;
                }
            }
            c += n;
        }
    }
    if constexpr (MODE == 1) {
        flush_sub(cur_sub);
    } else {
        float* o = out + out_start;
        if (use_atomic) {
            if (ta < nout) atomicAdd(o + ta, static_cast<float>(acc0));
            if (tb < nout) atomicAdd(o + tb, static_cast<float>(acc1));
        } else {
            if (ta < nout) o[ta] = static_cast<float>(acc0);
            if (tb < nout) o[tb] = static_cast<float>(acc1);
        }
    }
}

// float32 samples, channels summed in order: float64 for dedisperse (kernels.py:357-361
// accumulates in float64), float32 for subband (kernels.py:397-401 adds into the float32
// output); MODE as above.
template <int MODE>
__global__ __launch_bounds__(DD_THREADS) void k_gulp_f32(
    const float* __restrict__ in, const int32_t* __restrict__ delays,
    const int32_t* __restrict__ chan_to_sub, float* __restrict__ out, int64_t nchans, int64_t nsamps,
    int64_t nout, int64_t out_start, int64_t nsubs, int64_t out_size) {
    __shared__ float tile[DDF_ROWS * DDF_STRIDE];
    __shared__ int32_t sdel[DDF_CC];
    __shared__ int32_t ssub[DDF_CC];
    const int tid = threadIdx.x;
    const int sub_lane = tid & 15;
    const int row_grp = tid >> 4;
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * DD_T;
    const int64_t ta = t0 + tid;
    const int64_t tb = t0 + DD_THREADS + tid;
    using acc_t = std::conditional_t<MODE == 0, double, float>;
    acc_t acc0 = 0, acc1 = 0;
    int cur_sub = -1;

    auto flush_sub = [&](int sub) {
        if (sub >= 0) {
            const int64_t ia = ta * nsubs + sub, ib = tb * nsubs + sub;
            if (ta < nout && ia < out_size) atomicAdd(out + ia, static_cast<float>(acc0));
            if (tb < nout && ib < out_size) atomicAdd(out + ib, static_cast<float>(acc1));
        }
        acc0 = 0;
        acc1 = 0;
    };

    for (int64_t c = 0; c < nchans;) {
        const int ncand = static_cast<int>(min<int64_t>(DDF_CC, nchans - c));
        __syncthreads();
        if (tid < ncand) {
            sdel[tid] = delays[c + tid];
            if constexpr (MODE == 1) ssub[tid] = chan_to_sub[c + tid];
        }
        __syncthreads();
        int dmin, dmax;
        const int n = stage_extent(sdel, ncand, DDF_ROWS - DD_T, dmin, dmax);
        const int nrows = DD_T + (dmax - dmin);
        const int64_t rbase = t0 + dmin;
        if (sub_lane < n) {
            for (int r = row_grp; r < nrows; r += DD_THREADS / 16) {
                int64_t row = rbase + r;
                if (row < 0) row += nsamps;
                float v = 0.f;
                if (row >= 0 && row < nsamps) v = in[row * nchans + c + sub_lane];
                tile[r * DDF_STRIDE + sub_lane] = v;
            }
        }
        __syncthreads();
        for (int k = 0; k < n; ++k) {
            if constexpr (MODE == 1) {
                const int sub = ssub[k];
                if (sub != cur_sub) {
                    flush_sub(cur_sub);
                    cur_sub = sub;
                }
            }
            const int rel = sdel[k] - dmin;
            acc0 += static_cast<acc_t>(tile[(tid + rel) * DDF_STRIDE + k]);
            acc1 += static_cast<acc_t>(tile[(tid + DD_THREADS + rel) * DDF_STRIDE + k]);
        }
        c += n;
    }
    if constexpr (MODE == 1) {
        flush_sub(cur_sub);
    } else {
        if (ta < nout) out[out_start + ta] = static_cast<float>(acc0);
        if (tb < nout) out[out_start + tb] = static_cast<float>(acc1);
    }
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

int fill_f32(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return SPP_OK;
    const int64_t blocks = std::min<int64_t>(spp::ceil_div(n, 256), int64_t(spp::device_cu_count()) * 8);
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, st, p, n, 0.f);
    SPP_LAUNCHED("k_fill_f32");
    return SPP_OK;
}

// The dedisperse form with word-parallel adds (gulp_swar.hip), for rows of whole dwords.
// Channels are grouped as below, with the SWAR kernel's stage (a fixed run of dwords of each
// row) in place of DD_CC channels and its own time tile.
int launch_swar(const uint8_t* in, int nbits, int big, const int32_t* delays, float* out, int64_t nchans,
                int64_t nsamps, int64_t nout, int64_t out_start, hipStream_t st) {
    const int64_t row_bytes = nchans * nbits / 8;
    const bool compact = spp::gulp_swar_compact(nbits, nchans, nsamps - nout);
    const int64_t T = spp::gulp_swar_tile(nbits);
    const int64_t chunk = spp::gulp_swar_stage_channels(nbits);  // channels of one stage
    const int64_t bx = spp::ceil_div(nout, T);
    SPP_REQUIRE(bx <= 0x7fffffffLL, "gulp too long");
    const int64_t cus = spp::device_cu_count();
    const int64_t maxval = (int64_t(1) << nbits) - 1;
    const bool exact = nchans * maxval < (int64_t(1) << 24);
    int64_t groups = 1;
    if (exact)
        groups = std::max<int64_t>(1, std::min<int64_t>(spp::ceil_div(8 * cus, bx), spp::ceil_div(nchans, chunk)));
    // 1/2-bit counters are 16 bits wide: at most 21 760 channels per block
    if (nbits <= 2) groups = std::max<int64_t>(groups, spp::ceil_div(nchans, int64_t(21760)));
    int64_t cpg = spp::ceil_div(spp::ceil_div(nchans, groups), chunk) * chunk;
    int64_t ilv = 1, nchunk = 1;
    const int64_t line_ch = 1024 / nbits;  // channels in a 128-byte row line
    const int64_t line_spread = (nsamps - nout) * line_ch / nchans;
    const bool wide_groups = exact && cpg >= chunk && nchans >= line_ch && nchans % line_ch == 0;
    if (wide_groups && line_spread <= T) {
        cpg = chunk;
        ilv = line_ch / chunk;
        const int64_t hgroups = std::max<int64_t>(1, spp::ceil_div(groups, ilv));
        nchunk = spp::ceil_div(nchans, line_ch * hgroups);
        groups = ilv * spp::ceil_div(nchans, line_ch * nchunk);
    } else if (wide_groups) {
        cpg = chunk;
        groups = spp::ceil_div(nchans, chunk);
    } else {
        groups = spp::ceil_div(nchans, cpg);
    }
    const int use_atomic = groups > 1 ? 1 : 0;
    if (use_atomic) {
        const int rc = fill_f32(out + out_start, nout, st);
        if (rc) return rc;
    }
    return spp::launch_gulp_swar(in, nbits, big != 0, delays, out, nchans, nsamps, nout, out_start, row_bytes, cpg,
                                 groups, ilv, nchunk, use_atomic, compact, st);
}

template <int MODE>
int launch_int(const uint8_t* in, int nbits, int big, const int32_t* delays, const int32_t* cts,
               float* out, int64_t nchans, int64_t nsamps, int64_t nout, int64_t out_start,
               int64_t nsubs, int64_t out_size, hipStream_t st) {
    if (MODE == 0 && spp::gulp_swar_applies(in, nbits, nchans))
        return launch_swar(in, nbits, big, delays, out, nchans, nsamps, nout, out_start, st);
    const int64_t row_bytes = nchans * nbits / 8;
    const int64_t bx = spp::ceil_div(nout, DD_T);
    SPP_REQUIRE(bx <= 0x7fffffffLL, "gulp too long");
    // Split channels across blocks until the grid fills the device (integer sums are exact, so
    // the atomic combination is order-free).
    const int64_t cus = spp::device_cu_count();
    int64_t groups = 1;
    const int64_t maxval = (nbits == 8) ? 255 : ((1 << nbits) - 1);
    const bool exact = nchans * maxval < (int64_t(1) << 24);
    if (exact) {
        groups = std::max<int64_t>(1, std::min<int64_t>(spp::ceil_div(8 * cus, bx),
                                                        spp::ceil_div(nchans, int64_t(16))));
    }
    int64_t cpg = spp::ceil_div(nchans, groups);
    int64_t ilv = 1, nchunk = 1;
    // Dedisperse with wide groups uses groups of DD_CC channels (one stage each), numbered so
    // that the groups of a time tile run side by side on one XCD and read each packed row line
    // (128 bytes) from L2 once it is fetched. When the mean delay spread across one row line is
    // at most DD_T samples, the groups that share a line read mostly the same rows; then a block
    // takes nchunk chunks one line apart, and the ilv = line / chunk groups of a line run side
    // by side: the same single fetch with nchunk times fewer float atomics per output. Subband
    // otherwise keeps contiguous groups: its band partial sums are float atomics nsubs floats
    // apart, and splitting bands across groups multiplies them.
    // Subband takes the same interleave with chunks of whole bands (the mean band width rounded
    // up to DD_CC channels), so a band is still summed by one block and flushed once.
    const int64_t line_ch = 1024 / nbits;
    const int64_t line_spread = (nsamps - nout) * line_ch / nchans;
    const int64_t chunk = MODE == 0 ? DD_CC
                                    : spp::ceil_div(spp::ceil_div(nchans, std::max<int64_t>(1, nsubs)),
                                                    int64_t(DD_CC)) * DD_CC;
    const bool wide_groups = exact && cpg >= DD_CC && nchans >= line_ch;
    if (wide_groups && line_spread <= DD_T && line_ch % chunk == 0) {
        cpg = chunk;
        ilv = line_ch / chunk;
        const int64_t hgroups = std::max<int64_t>(1, spp::ceil_div(groups, ilv));
        nchunk = spp::ceil_div(nchans, line_ch * hgroups);
        groups = ilv * spp::ceil_div(nchans, line_ch * nchunk);
    } else if (MODE == 0 && wide_groups) {
        cpg = DD_CC;
        groups = spp::ceil_div(nchans, int64_t(DD_CC));
    } else {
        // Group starts on multiples of 64 channels keep full stages dword-aligned (a power-of-two
        // unit count per stage); narrow groups on multiples of 8.
        cpg = cpg > 64 ? spp::ceil_div(cpg, 64) * 64 : spp::ceil_div(cpg, 8) * 8;
        groups = spp::ceil_div(nchans, cpg);
    }
    const int use_atomic = (MODE == 0 && groups > 1) ? 1 : 0;
    if (use_atomic) {
        int rc = fill_f32(out + out_start, nout, st);
        if (rc) return rc;
    }
    const int64_t nblocks = bx * groups;
    SPP_REQUIRE(nblocks <= 0x7ffffff0LL, "gulp too long");
    dim3 grid(static_cast<unsigned>(spp::ceil_div(nblocks, int64_t(8)) * 8));
    const bool wide = row_bytes % 4 == 0 && reinterpret_cast<uintptr_t>(in) % 4 == 0;
    const bool narrow = cpg <= 16;
    // 64-channel tiles of DD_ROWS_LOW rows (4 blocks per CU instead of 3) when twice the mean
    // delay spread across DD_CC channels fits them; wider spreads split into more stages there.
    const bool low = !narrow && (nsamps - nout) * 2 * DD_CC <= int64_t(DD_ROWS_LOW - DD_T) * nchans;
#define SPP_GULP_ONE(NB, BG, W, CX, R)                                                              \
    hipLaunchKernelGGL((k_gulp_int<NB, BG, MODE, W, CX, R>), grid, dim3(DD_THREADS), 0, st, in, delays, cts, \
                       out, nchans, nsamps, nout, out_start, row_bytes, cpg, use_atomic, nsubs, out_size, groups, \
                       nblocks, ilv, nchunk)
#define SPP_GULP_LAUNCH(NB, BG)                            \
    if (wide && narrow)                                    \
        SPP_GULP_ONE(NB, BG, true, 16, DD_ROWS);           \
    else if (wide && low)                                  \
        SPP_GULP_ONE(NB, BG, true, DD_CC, DD_ROWS_LOW);    \
    else if (wide)                                         \
        SPP_GULP_ONE(NB, BG, true, DD_CC, DD_ROWS);        \
    else if (narrow)                                       \
        SPP_GULP_ONE(NB, BG, false, 16, DD_ROWS);          \
    else if (low)                                          \
        SPP_GULP_ONE(NB, BG, false, DD_CC, DD_ROWS_LOW);   \
    else                                                   \
        SPP_GULP_ONE(NB, BG, false, DD_CC, DD_ROWS)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_GULP_LAUNCH(1, false); break;
        case 3: SPP_GULP_LAUNCH(1, true); break;
        case 4: SPP_GULP_LAUNCH(2, false); break;
        case 5: SPP_GULP_LAUNCH(2, true); break;
        case 8: SPP_GULP_LAUNCH(4, false); break;
        case 9: SPP_GULP_LAUNCH(4, true); break;
        default: SPP_GULP_LAUNCH(8, false); break;
    }
#undef SPP_GULP_LAUNCH
#undef SPP_GULP_ONE
    SPP_LAUNCHED("k_gulp_int");
    return SPP_OK;
}

int check_common(const void* in, int nbits, int64_t nchans, int64_t nsamps, const int32_t* delays) {
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 32,
                "nbits must be 1, 2, 4, 8 or 32, got %d", nbits);
    SPP_REQUIRE(nchans > 0, "nchans must be > 0");
    SPP_REQUIRE(nsamps > 0, "nsamps must be > 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(in != nullptr && delays != nullptr, "null buffer");
    return SPP_OK;
}

}  // namespace

extern "C" int spp_dedisperse(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                              const int32_t* delays, int64_t max_delay, float* out,
                              int64_t out_start, void* stream) {
    spp::clear_error();
    int rc = check_common(in, nbits, nchans, nsamps, delays);
    if (rc) return rc;
    SPP_REQUIRE(out != nullptr, "null buffer");
    SPP_REQUIRE(out_start >= 0, "out_start must be >= 0");
    SPP_REQUIRE(max_delay >= 0, "max_delay must be >= 0");
    hipStream_t st = spp::as_stream(stream);
    const int64_t nout = nsamps - max_delay;
    if (nout <= 0) return SPP_OK;  // the reference loop body never runs
    if (nbits == 32) {
        const int64_t bx = spp::ceil_div(nout, DD_T);
        hipLaunchKernelGGL((k_gulp_f32<0>), dim3(static_cast<unsigned>(bx)), dim3(DD_THREADS), 0, st,
                           static_cast<const float*>(in), delays, nullptr, out, nchans, nsamps, nout,
                           out_start, int64_t(0), int64_t(0));
        SPP_LAUNCHED("k_gulp_f32");
        return SPP_OK;
    }
    return launch_int<0>(static_cast<const uint8_t*>(in), nbits, big, delays, nullptr, out, nchans,
                         nsamps, nout, out_start, 0, 0, st);
}

extern "C" int spp_subband(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                           const int32_t* delays, const int32_t* chan_to_sub, int64_t maxdelay,
                           int64_t nsubs, float* out, int64_t out_size, void* stream) {
    spp::clear_error();
    int rc = check_common(in, nbits, nchans, nsamps, delays);
    if (rc) return rc;
    SPP_REQUIRE(out != nullptr && chan_to_sub != nullptr, "null buffer");
    SPP_REQUIRE(nsubs > 0, "nsubs must be > 0");
    SPP_REQUIRE(out_size >= 0, "out_size must be >= 0");
    SPP_REQUIRE(maxdelay >= 0, "maxdelay must be >= 0");
    hipStream_t st = spp::as_stream(stream);
    const int64_t nout = nsamps - maxdelay;
    if (nout <= 0) return SPP_OK;
    rc = fill_f32(out, std::min<int64_t>(out_size, nout * nsubs), st);
    if (rc) return rc;
    if (nbits == 32) {
        const int64_t bx = spp::ceil_div(nout, DD_T);
        hipLaunchKernelGGL((k_gulp_f32<1>), dim3(static_cast<unsigned>(bx)), dim3(DD_THREADS), 0, st,
                           static_cast<const float*>(in), delays, chan_to_sub, out, nchans, nsamps,
                           nout, int64_t(0), nsubs, out_size);
        SPP_LAUNCHED("k_gulp_f32");
        return SPP_OK;
    }
    return launch_int<1>(static_cast<const uint8_t*>(in), nbits, big, delays, chan_to_sub, out,
                         nchans, nsamps, nout, 0, nsubs, out_size, st);
}
