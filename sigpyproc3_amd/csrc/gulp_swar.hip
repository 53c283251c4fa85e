// One-DM dedispersion of a packed 1/2/4/8-bit gulp with word-parallel adds: the dedisperse
// form of kernels.dedisperse (core/kernels.py:340-362) for integer samples, when every
// packed row is whole dwords.
//
// A block owns T consecutive outputs and a group of channels, taken in stages of GS_S packed
// dword columns (16 bytes of each row; F = 32/nbits channels per dword). Staging turns rows
// into channel-major series: a thread loads F consecutive rows of one dword column (an F x F
// block of samples), transposes it in registers by delta swaps (log2 F rounds of shift,
// mask and xor between word pairs), and stores word p -- F consecutive samples of the
// column's channel p -- into that channel's series in LDS. Summing, a thread owns F
// consecutive outputs: for each channel it reads the two series dwords around its delayed
// start, funnel-shifts them (v_alignbit) to the F samples it needs and adds them into packed
// counters of twice the sample width, widened to 8, 16 and 32 bits before they can
// overflow. Integer sums are exact, so channel groups combine by float atomics, order-free.
//
// Work per sample staged: 1/F of a dword load and of an LDS store and about 2 log2(F)/F
// transpose ops. Per channel and F outputs: one 8-byte LDS read, one funnel shift and five
// adds and masks. For 2-bit data (F = 16) that is about 0.5 ops per sample staged and 0.45
// per output-channel. Time tiles are 2048 outputs (1/2-bit) or 1024 (4/8-bit): long enough
// that one channel group per 16-byte row chunk takes a whole line of work, so the groups of
// a 128-byte row line can run side by side (the line interleave) and share it in L2.
//
// A stage's channels whose delays span more than SPREAD samples are split into passes
// (consecutive columns, or runs of one column's channels), each staging its own row window.
#include "spp_gulp_swar.hpp"

#include "spp_common.hpp"

#include <algorithm>

namespace {

constexpr int GS_THREADS = 256;
constexpr int GS_S = 4;     // dword columns per stage (16 bytes of each row)
constexpr int GS_LOGS = 2;  // log2(GS_S)
constexpr int GS_RMAX = 2560;  // rows one pass stages at most (the compact form: fewer, below)
// Word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format, no swizzle.
constexpr int kBufferFlags = 0x00020000;

// Row pitch of the series (dwords per channel): at least nrb, and with f * pitch = 16
// (mod 64) where that exists, so that the four columns a wave stores land 16 banks apart.
constexpr int series_pitch(int f, int nrb) {
    for (int p = 0; p < 64; ++p)
        if ((f * (nrb + p)) % 64 == 16) return nrb + p;
    return nrb + 1;
}

// COMPACT (4/8-bit data): a series buffer of half the rows, about 21 KB, so that twice as many
// blocks share a CU (these kernels are bound by LDS size, not registers) when the stages' delay
// spreads are small. 1/2-bit blocks are register-bound: more LDS room buys them nothing.
template <int NBITS, bool COMPACT = false>
struct Geo {
    static constexpr int F = 32 / NBITS;                                    // samples per dword
    static constexpr int T = (NBITS == 8 || NBITS == 4) ? 1024 : 2048;     // outputs per block
    static constexpr int RMAX = COMPACT ? 1280 : GS_RMAX;
    static constexpr int NG = T / F > GS_THREADS ? T / F / GS_THREADS : 1;  // output groups per thread
    static constexpr int G = T / F / NG;                                    // threads per partition
    static constexpr int H = GS_THREADS / G;                                // channel partitions
    static constexpr int NRB = RMAX / F;                                    // series dwords used
    static constexpr int PITCH = series_pitch(F, NRB);
    static constexpr int SPREAD = RMAX - T - 2 * F;                         // per pass
    static constexpr int CH = GS_S * F;                                     // channels per stage
    static_assert(G * H == GS_THREADS && G >= 64 && G * NG * F == T, "a wave's lanes share one channel partition");
    static_assert(!COMPACT || NBITS >= 4, "compact form for byte-wide and nibble-wide samples only");
};

constexpr uint32_t swap_mask(int s) {
    return s == 16 ? 0x0000ffffu : s == 8 ? 0x00ff00ffu : s == 4 ? 0x0f0f0f0fu : s == 2 ? 0x33333333u : 0x55555555u;
}

// Transpose of an F x F block of NBITS-bit elements: element c of word r (bits NBITS*c ..)
// moves to element r of word c.
template <int NBITS>
__device__ __forceinline__ void transpose_block(uint32_t (&w)[32 / NBITS]) {
    constexpr int F = 32 / NBITS;
    constexpr int LOGF = NBITS == 1 ? 5 : (NBITS == 2 ? 4 : (NBITS == 4 ? 3 : 2));
#pragma unroll
    for (int lv = 0; lv < LOGF; ++lv) {
        const int je = F >> (lv + 1);
        const int s = je * NBITS;
        const uint32_t m = swap_mask(s);
#pragma unroll
        for (int r = 0; r < F; ++r) {
            if (r & je) continue;
            const uint32_t t = ((w[r] >> s) ^ w[r + je]) & m;
            w[r] ^= t << s;
            w[r + je] ^= t;
        }
    }
}

// Channel (within its dword) of the field at bits NBITS*p: byte p / E, and in the byte the
// first channel sits in the top bits when BIG (E = 8 / NBITS fields per byte).
template <int NBITS, bool BIG>
__device__ __forceinline__ constexpr int chan_of(int p) {
    constexpr int E = NBITS >= 8 ? 1 : 8 / NBITS;
    return BIG ? (p / E) * E + (E - 1 - p % E) : p;
}

// Packed counters for F outputs. add(v): v holds the F samples (NBITS each) of one channel.
// The call pattern (and so the flush counters) is the same in every lane of a wave.
template <int NBITS>
struct SwarAcc;

template <>
struct SwarAcc<1> {  // nibbles (15 channels), bytes (17 flushes), 16-bit fields
    uint32_t n[4] = {0, 0, 0, 0}, b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t h[16] = {};
    int k0 = 0, k1 = 0;
    __device__ __forceinline__ void add(uint32_t v) {
#pragma unroll
        for (int j = 0; j < 4; ++j) n[j] += (v >> j) & 0x11111111u;  // nibble k: sample 4k + j
        if (++k0 == 15) flush0();
    }
    __device__ __forceinline__ void flush0() {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b[2 * j] += n[j] & 0x0f0f0f0fu;             // byte m: sample 8m + j
            b[2 * j + 1] += (n[j] >> 4) & 0x0f0f0f0fu;  // byte m: sample 8m + 4 + j
            n[j] = 0;
        }
        k0 = 0;
        if (++k1 == 17) flush1();
    }
    __device__ __forceinline__ void flush1() {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            h[2 * i] += b[i] & 0x00ff00ffu;             // field f: byte 2f
            h[2 * i + 1] += (b[i] >> 8) & 0x00ff00ffu;  // field f: byte 2f + 1
            b[i] = 0;
        }
        k1 = 0;
    }
    __device__ __forceinline__ void finish(uint32_t (&o)[32]) {
        flush0();
        flush1();
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int m = j >> 3, i = 2 * (j & 3) + ((j >> 2) & 1);
            o[j] = (h[2 * i + (m & 1)] >> (16 * (m >> 1))) & 0xffffu;
        }
    }
};

template <>
struct SwarAcc<2> {  // nibbles (5 channels), bytes (16 flushes), 16-bit fields
    uint32_t ne = 0, no = 0, b[4] = {0, 0, 0, 0};
    uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int k0 = 0, k1 = 0;
    __device__ __forceinline__ void add(uint32_t v) {
        ne += v & 0x33333333u;         // nibble k: sample 2k
        no += (v >> 2) & 0x33333333u;  // nibble k: sample 2k + 1
        if (++k0 == 5) flush0();
    }
    __device__ __forceinline__ void flush0() {
        b[0] += ne & 0x0f0f0f0fu;         // byte m: sample 4m
        b[1] += (ne >> 4) & 0x0f0f0f0fu;  // 4m + 2
        b[2] += no & 0x0f0f0f0fu;         // 4m + 1
        b[3] += (no >> 4) & 0x0f0f0f0fu;  // 4m + 3
        ne = no = 0;
        k0 = 0;
        if (++k1 == 16) flush1();
    }
    __device__ __forceinline__ void flush1() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h[2 * i] += b[i] & 0x00ff00ffu;
            h[2 * i + 1] += (b[i] >> 8) & 0x00ff00ffu;
            b[i] = 0;
        }
        k1 = 0;
    }
    __device__ __forceinline__ void finish(uint32_t (&o)[16]) {
        flush0();
        flush1();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int m = j >> 2, r = j & 3;
            const int i = r == 0 ? 0 : (r == 1 ? 2 : (r == 2 ? 1 : 3));
            o[j] = (h[2 * i + (m & 1)] >> (16 * (m >> 1))) & 0xffffu;
        }
    }
};

template <>
struct SwarAcc<4> {  // bytes (17 channels), 16-bit fields (257 flushes), 32-bit
    uint32_t be = 0, bo = 0, h[4] = {0, 0, 0, 0};
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int k0 = 0, k1 = 0;
    __device__ __forceinline__ void add(uint32_t v) {
        be += v & 0x0f0f0f0fu;         // byte m: sample 2m
        bo += (v >> 4) & 0x0f0f0f0fu;  // 2m + 1
        if (++k0 == 17) flush0();
    }
    __device__ __forceinline__ void flush0() {
        h[0] += be & 0x00ff00ffu;         // field f: sample 4f
        h[1] += (be >> 8) & 0x00ff00ffu;  // 4f + 2
        h[2] += bo & 0x00ff00ffu;         // 4f + 1
        h[3] += (bo >> 8) & 0x00ff00ffu;  // 4f + 3
        be = bo = 0;
        k0 = 0;
        if (++k1 == 257) flush1();
    }
    __device__ __forceinline__ void flush1() {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            w[4 * f + 0] += (h[0] >> (16 * f)) & 0xffffu;
            w[4 * f + 2] += (h[1] >> (16 * f)) & 0xffffu;
            w[4 * f + 1] += (h[2] >> (16 * f)) & 0xffffu;
            w[4 * f + 3] += (h[3] >> (16 * f)) & 0xffffu;
        }
        h[0] = h[1] = h[2] = h[3] = 0;
        k1 = 0;
    }
    __device__ __forceinline__ void finish(uint32_t (&o)[8]) {
        flush0();
        flush1();
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = w[j];
    }
};

template <>
struct SwarAcc<8> {  // 16-bit fields (257 channels), 32-bit
    uint32_t he = 0, ho = 0, w[4] = {0, 0, 0, 0};
    int k0 = 0;
    __device__ __forceinline__ void add(uint32_t v) {
        he += v & 0x00ff00ffu;         // field f: sample 2f
        ho += (v >> 8) & 0x00ff00ffu;  // 2f + 1
        if (++k0 == 257) flush0();
    }
    __device__ __forceinline__ void flush0() {
        w[0] += he & 0xffffu;
        w[2] += he >> 16;
        w[1] += ho & 0xffffu;
        w[3] += ho >> 16;
        he = ho = 0;
        k0 = 0;
    }
    __device__ __forceinline__ void finish(uint32_t (&o)[4]) {
        flush0();
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = w[j];
    }
};

template <int NBITS, bool BIG, bool COMPACT>
__global__ __launch_bounds__(GS_THREADS, 3) void k_gulp_swar(
    const uint8_t* __restrict__ in, const int32_t* __restrict__ delays, float* __restrict__ out, int64_t nchans,
    int64_t nsamps, int64_t nout, int64_t out_start, int64_t row_bytes, int64_t chans_per_group, int64_t ngroups,
    int64_t nblocks, int64_t ilv, int64_t nchunk, int use_atomic) {
    using Gm = Geo<NBITS, COMPACT>;
    constexpr int F = Gm::F, T = Gm::T, G = Gm::G, H = Gm::H, NG = Gm::NG, PITCH = Gm::PITCH, CH = Gm::CH;
    constexpr int UNITS = F >= 32 ? 1 : 2;  // staging units in flight per thread
    __shared__ uint32_t series[CH * PITCH];
    __shared__ __align__(16) int32_t sdel[CH];
    __shared__ int2 cinfo[CH];  // per stage channel: series index of its first dword, funnel shift
    __shared__ int32_t cmin[GS_S], cmax[GS_S];
    // XCD-aware order, as the byte-staged kernel: the groups of one time tile are consecutive
    // logical blocks on one XCD, and share the packed row lines in its L2.
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (logical >= nblocks) return;
    const int tid = threadIdx.x;
    const int g = tid % G, h = tid / G;
    const int64_t t0 = (logical / ngroups) * T;
    const int64_t grp = logical % ngroups;
    const int64_t gl = grp % ilv, gh = grp / ilv;
    SwarAcc<NBITS> acc[NG];  // output groups g + G i, i < NG

    // The block's stages in order: chunk m of chans_per_group channels (a multiple of CH),
    // CH channels at a time; (m, j) is stage j of chunk m.
    auto stage_at = [&](int64_t m, int64_t j, int64_t& c, int& nch) {
        if (m >= nchunk) return false;
        const int64_t cb = ((gh * nchunk + m) * ilv + gl) * chans_per_group;
        c = cb + j * CH;
        const int64_t ce = min(nchans, cb + chans_per_group);
        if (c >= ce) return false;  // past nchans: so are all later chunks
        nch = static_cast<int>(min<int64_t>(CH, ce - c));  // a multiple of F
        return true;
    };
    const int64_t spc = (chans_per_group + CH - 1) / CH;  // stages per chunk

    // One pass: columns [q0, q1) staged for the delay window [dmin, dmax], channels [c0, c1)
    // summed.
    auto run_pass = [&](int64_t c, int q0, int q1, int c0, int c1, int dmin, int dmax) {
        if (tid >= c0 && tid < c1) {
            const int rel = sdel[tid] - dmin;
            cinfo[tid] = make_int2(tid * PITCH + rel / F, NBITS * (rel % F));
        }
        // Staging: units (column q, row block b) -> rows rbase + F b + k, k < F.
        const int nqp = q1 - q0;
        const int nrb = T / F + (dmax - dmin + F - 1) / F + 1;
        const int units = nqp * nrb;
        const int64_t rbase = t0 + dmin;
        const int64_t col0 = (c / F + q0) * 4;  // byte offset of the pass's first column
        // The pass's rows as one buffer when they all lie in the stream and span under 2^31
        // bytes (almost always): a unit's F loads then share one offset register and step by
        // row_bytes in the scalar offset.
        const int64_t span = int64_t(F) * nrb * row_bytes;
        const bool inside = rbase >= 0 && rbase + int64_t(F) * nrb <= nsamps && span < (int64_t(1) << 31);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint8_t*>(in) + (inside ? rbase * row_bytes + col0 : 0), 0,
            inside ? static_cast<int>(span) : 0, kBufferFlags);
        // unit u -> (column u % nqp, row block u / nqp); shifts when the pass has all GS_S columns
        const bool full_width = nqp == GS_S;
        auto unit_col = [&](int u) { return full_width ? (u & (GS_S - 1)) : u % nqp; };
        auto unit_blk = [&](int u) { return full_width ? (u >> GS_LOGS) : u / nqp; };
        auto load_unit = [&](int u, uint32_t (&w)[F]) {
            const int qq = unit_col(u), b = unit_blk(u);
            if (inside) {
                const uint32_t off = static_cast<uint32_t>(F * b) * static_cast<uint32_t>(row_bytes) + 4u * qq;
#pragma unroll
                for (int k = 0; k < F; ++k)
                    w[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, off, k * static_cast<int>(row_bytes), 0);
                return;
            }
            const int64_t r0 = rbase + int64_t(F) * b;
            {
#pragma unroll
                for (int k = 0; k < F; ++k) {
                    int64_t row = r0 + k;
                    if (row < 0) row += nsamps;  // numba wraparound of a negative index
                    w[k] = (row >= 0 && row < nsamps)
                               ? *reinterpret_cast<const uint32_t*>(in + row * row_bytes + col0 + 4 * qq)
                               : 0u;
                }
            }
        };
        auto store_unit = [&](int u, uint32_t (&w)[F]) {
            transpose_block<NBITS>(w);
            const int cq = (q0 + unit_col(u)) * F, b = unit_blk(u);
#pragma unroll
            for (int p = 0; p < F; ++p) series[(cq + chan_of<NBITS, BIG>(p)) * PITCH + b] = w[p];
        };
        for (int u = tid; u < units; u += UNITS * GS_THREADS) {
            uint32_t w[UNITS][F];
#pragma unroll
            for (int i = 0; i < UNITS; ++i)
                if (u + i * GS_THREADS < units) load_unit(u + i * GS_THREADS, w[i]);
#pragma unroll
            for (int i = 0; i < UNITS; ++i)
                if (u + i * GS_THREADS < units) store_unit(u + i * GS_THREADS, w[i]);
        }
        __syncthreads();
        // Summing: this thread's outputs t0 + F (g + G i) + j over its share of the pass's
        // channels (wave-uniform: a wave's lanes share h), four channels' reads in flight at
        // a time.
        const int len = c1 - c0;
        const int k0 = c0 + (len * h) / H, k1 = c0 + (len * (h + 1)) / H;
        int k = k0;
        for (; k + 4 <= k1; k += 4) {
            int2 ci[4];
            uint32_t lo[4][NG], hi[4][NG];
#pragma unroll
            for (int j = 0; j < 4; ++j) ci[j] = cinfo[k + j];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < NG; ++i) {
                    lo[j][i] = series[ci[j].x + g + G * i];
                    hi[j][i] = series[ci[j].x + g + G * i + 1];
                }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < NG; ++i)
                    acc[i].add(__builtin_amdgcn_alignbit(hi[j][i], lo[j][i], static_cast<uint32_t>(ci[j].y)));
        }
        for (; k < k1; ++k) {
            const int2 ci = cinfo[k];
#pragma unroll
            for (int i = 0; i < NG; ++i)
                acc[i].add(__builtin_amdgcn_alignbit(series[ci.x + g + G * i + 1], series[ci.x + g + G * i],
                                                     static_cast<uint32_t>(ci.y)));
        }
    };

    int64_t c = 0, sm = 0, sj = 0;
    int nch = 0;
    bool have = stage_at(0, 0, c, nch);
    int32_t dpre = (have && tid < nch) ? delays[c + tid] : 0;
    while (have) {
        __syncthreads();  // the previous stage's tables and series are no longer read
        if (tid < nch) sdel[tid] = dpre;
        if (++sj == spc) {
            sj = 0;
            ++sm;
        }
        int64_t cn = 0;
        int nchn = 0;
        const bool haven = stage_at(sm, sj, cn, nchn);
        if (haven && tid < nchn) dpre = delays[cn + tid];  // the next stage's delays load meanwhile
        __syncthreads();
        const int nq = nch / F;
        if (tid < nq) {
            const int4* d4 = reinterpret_cast<const int4*>(sdel + tid * F);
            int mn = d4[0].x, mx = mn;
#pragma unroll
            for (int i = 0; i < F / 4; ++i) {
                const int4 v = d4[i];
                mn = min(min(mn, v.x), min(v.y, min(v.z, v.w)));
                mx = max(max(mx, v.x), max(v.y, max(v.z, v.w)));
            }
            cmin[tid] = mn;
            cmax[tid] = mx;
        }
        __syncthreads();
        // Passes, planned alike in every thread: runs of whole columns whose delays span at
        // most SPREAD; a column that alone spans more goes in runs of its channels.
        int q = 0, kk = 0;
        bool first = true;
        while (q < nq) {
            if (!first) __syncthreads();  // the previous pass's tables and series are no longer read
            first = false;
            if (cmax[q] - cmin[q] > Gm::SPREAD) {
                int mn = sdel[q * F + kk], mx = mn, e = kk + 1;
                while (e < F) {
                    const int d = sdel[q * F + e];
                    if (max(mx, d) - min(mn, d) > Gm::SPREAD) break;
                    mn = min(mn, d);
                    mx = max(mx, d);
                    ++e;
                }
                run_pass(c, q, q + 1, q * F + kk, q * F + e, mn, mx);
                kk = e;
                if (kk == F) {
                    kk = 0;
                    ++q;
                }
                continue;
            }
            int mn = cmin[q], mx = cmax[q], e = q + 1;
            while (e < nq && cmax[e] - cmin[e] <= Gm::SPREAD && max(mx, cmax[e]) - min(mn, cmin[e]) <= Gm::SPREAD) {
                mn = min(mn, cmin[e]);
                mx = max(mx, cmax[e]);
                ++e;
            }
            run_pass(c, q, e, q * F, e * F, mn, mx);
            q = e;
        }
        c = cn;
        nch = nchn;
        have = haven;
    }
    // The H partitions' sums of each output, then one store (or float atomic) per output.
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        uint32_t o[F];
        acc[i].finish(o);
#pragma unroll
        for (int j = 0; j < F; ++j) series[h * T + F * (g + G * i) + j] = o[j];
    }
    __syncthreads();
    float* dst = out + out_start;
    for (int t = tid; t < T; t += GS_THREADS) {
        uint32_t s = 0;
#pragma unroll
        for (int hh = 0; hh < H; ++hh) s += series[hh * T + t];
        const int64_t ti = t0 + t;
        if (ti < nout) {
            if (use_atomic)
                atomicAdd(dst + ti, static_cast<float>(s));
            else
                dst[ti] = static_cast<float>(s);
        }
    }
}

// Largest channel count per block for which the 16-bit counters of 1- and 2-bit data cannot
// overflow (4- and 8-bit counters widen to 32 bits).
int64_t swar_max_channels(int nbits) {
    return nbits == 1 ? 65535 : (nbits == 2 ? 21760 : (int64_t(1) << 40));
}

}  // namespace

namespace spp {

int64_t gulp_swar_tile(int nbits) {
    return nbits == 8 ? Geo<8>::T : (nbits == 4 ? Geo<4>::T : (nbits == 2 ? Geo<2>::T : Geo<1>::T));
}

bool gulp_swar_compact(int nbits, int64_t nchans, int64_t max_delay) {
    if (nbits < 4 || nchans <= 0) return false;
    const int64_t spread = nbits == 8 ? Geo<8, true>::SPREAD : Geo<4, true>::SPREAD;
    // twice the mean delay spread of a stage's channels (the widest stages of a dispersion
    // sweep run 1.4-1.6 times the mean)
    return 2 * max_delay * gulp_swar_stage_channels(nbits) <= spread * nchans;
}

int64_t gulp_swar_stage_channels(int nbits) { return int64_t(GS_S) * (32 / nbits); }

bool gulp_swar_applies(const void* in, int nbits, int64_t nchans) {
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8) return false;
    return (nchans * nbits) % 32 == 0 && reinterpret_cast<uintptr_t>(in) % 4 == 0;
}

int launch_gulp_swar(const uint8_t* in, int nbits, bool big, const int32_t* delays, float* out, int64_t nchans,
                     int64_t nsamps, int64_t nout, int64_t out_start, int64_t row_bytes, int64_t chans_per_group,
                     int64_t groups, int64_t ilv, int64_t nchunk, int use_atomic, bool compact, hipStream_t st) {
    SPP_REQUIRE(chans_per_group * nchunk <= swar_max_channels(nbits), "too many channels per block");
    SPP_REQUIRE(!compact || nbits >= 4, "the compact form is for 4- and 8-bit data");
    const int64_t bx = ceil_div(nout, gulp_swar_tile(nbits));
    const int64_t nblocks = bx * groups;
    SPP_REQUIRE(nblocks <= 0x7ffffff0LL, "gulp too long");
    const dim3 grid(static_cast<unsigned>(ceil_div(nblocks, int64_t(8)) * 8));
#define SPP_SWAR(NB, BG, CP)                                                                                   \
    hipLaunchKernelGGL((k_gulp_swar<NB, BG, CP>), grid, dim3(GS_THREADS), 0, st, in, delays, out, nchans, nsamps, \
                       nout, out_start, row_bytes, chans_per_group, groups, nblocks, ilv, nchunk, use_atomic)
    switch (nbits * 4 + (big && nbits < 8 ? 2 : 0) + (compact ? 1 : 0)) {  // bit order: sub-byte samples only
        case 4: SPP_SWAR(1, false, false); break;
        case 6: SPP_SWAR(1, true, false); break;
        case 8: SPP_SWAR(2, false, false); break;
        case 10: SPP_SWAR(2, true, false); break;
        case 16: SPP_SWAR(4, false, false); break;
        case 17: SPP_SWAR(4, false, true); break;
        case 18: SPP_SWAR(4, true, false); break;
        case 19: SPP_SWAR(4, true, true); break;
        case 33: SPP_SWAR(8, false, true); break;
        default: SPP_SWAR(8, false, false); break;
    }
#undef SPP_SWAR
    SPP_LAUNCHED("k_gulp_swar");
    return SPP_OK;
}

}  // namespace spp
