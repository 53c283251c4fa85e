// DM-trial sweep: plane[i][t] = sum_c x_c[t + d_i(c)] for every DM row i of a delay
// table, over a whole stream resident in HBM.
//
// Stage 1 (k_build_g) turns the packed time-major stream [nsamps][nchans] into a
// channel-major, time-interleaved word array G[c][j] (uint32) holding IL samples
// x_c[j + k*S] (k < IL, S = 256) in fields of 32/IL bits, zero past the end:
//   1- and 2-bit data: IL = 8 nibble fields; 4- and 8-bit data: IL = 4 byte fields.
// A delay shift d of channel c moves every field of a word together, so an output word
// (the IL samples t + k*S) is the sum over channels of the aligned words G[c][t + d_i(c)].
//
// Stage 2 (k_sweep): a block owns a DM tile (<= 16 consecutive DM rows whose
// per-channel delay spread is <= SW_SPREAD_MAX, formed by the host planner) and IL*256
// output samples. Channels are processed in stages of CC: wave w stages channels
// w, w + WAVES, ... of the next stage, reading the windows G[c][jb .. jb + 256 + spread]
// (jb = t0 + min delay of the tile, rounded down to 4 words) with raw buffer loads whose
// range is the channel's valid words (past them the hardware returns 0). The loads carry the
// LDS flag: they write the other half of a double-buffered LDS area directly while the
// current stage is summed (one vmcnt wait and one barrier per stage). Each wave owns K DM rows; per channel and DM row a lane
// reads 4 LDS words (lanes read consecutive words: bank-conflict free) at a byte offset
// the planner's relative delays give (scalar loads, one stage's row at a time), and adds
// them into SWAR accumulators:
//   nibble fields (1/2-bit): groups of GN channels (GN * maxval <= 15: 5 for 2-bit) summed
//                  with v_add3; per pair of groups the odd fields are masked into byte lanes
//                  and the raw words summed (even = raw - 16 * odd, recovered when byte
//                  lanes are widened to 16-bit lanes every SB stages);
//   byte fields (4-bit): summed over the stage, then widened to 16-bit lanes;
//   8-bit data: odd bytes (v_perm) go straight to 16-bit lanes and the raw words are summed
//                  as 32-bit integers; even = raw - 256 * odd when the lanes are read;
//   32-bit lanes only when nchans * maxval > 65535.
// Integer sums are exact, so the float32 output is bit-exact to the reference's float64 sums.
//
// Work: nchans adds per DM-trial sample. LDS bytes read per sample-add: 4/IL (0.5 for
// nibble fields, 1 for byte fields). The sweep is bound by LDS reads and VALU issue, not HBM.
#include "spp_common.hpp"

#include <algorithm>
#include <new>
#include <vector>

namespace {

typedef __attribute__((address_space(3))) void lds_void_t;  // an LDS address for the LDS-load builtin

constexpr int SW_S = 256;                // interleave stride = words per wave row
constexpr int SW_DT = 16;                // DM rows per tile
constexpr int SW_LW = 512;               // LDS words per staged channel (2 KiB: a multiple of the
                                         // 256-byte unit of ds_read2st64 offsets)
constexpr int SW_SPREAD_MAX = SW_LW - SW_S - 4;  // per-channel delay spread of a DM tile
constexpr int B_TJ = 128;                // words per build block (time)
constexpr int B_THREADS = 512;           // threads per build block
constexpr int G_STRIDE = 68;             // bytes per staged row in the build kernel
static_assert(SW_S == 2 * B_TJ, "stage-1 blocks two apart share staged rows");

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

template <int NBITS>
struct Fields {
    static constexpr int IL = NBITS <= 2 ? 8 : 4;          // samples per word
    static constexpr int FB = 32 / IL;                     // bits per field
    static constexpr uint32_t MAXV = NBITS == 8 ? 255u : ((1u << NBITS) - 1u);
};

// Stage 1. Block (bx, by): words [bx*B_TJ, bx*B_TJ + B_TJ) of the channels in row bytes
// [by*64, by*64 + 64) (64 * 8/NBITS channels). The rows j0 + k*S + i (k < IL, i < B_TJ) are
// staged in LDS. Blocks are numbered with by fastest and consecutive numbers on one XCD, so
// the column blocks of a row range (which share its 128-byte lines) and the blocks bx,
// bx + 2, ..., bx + 2(IL - 1) (which stage the same rows, S being 2 * B_TJ) run close together
// and re-read those rows from that XCD's L2. Thread (quad q, byte column bc) then builds
// words j0 + 4q .. 4q+3 of every channel packed in byte column bc from one LDS byte per
// (row, word) and stores them as one 16-byte write per channel (a wave writes 512 contiguous
// bytes per channel). Only the first glen words of each channel row are written (the sweep
// reads no further).
template <int NBITS, bool BIG, bool ALIGNED>
__global__ __launch_bounds__(B_THREADS) void k_build_g(const uint8_t* __restrict__ in, int64_t nsamps,
                                                       int64_t nchans, uint32_t* __restrict__ G,
                                                       int64_t gpitch, int64_t glen) {
    using F = Fields<NBITS>;
    constexpr int IL = F::IL;
    constexpr int PER = (NBITS < 8) ? 8 / NBITS : 1;     // channels per byte
    constexpr int ROWS = IL * B_TJ;
    // Staged row r sits at position (r % 4) * RQ + r / 4: the rows 4qd + u a build step reads
    // for one u are then consecutive (one bank each), and the 4 rows a staging wave writes
    // fall on different banks.
    constexpr int RQ = ROWS / 4 + 1;
    constexpr int NE = ROWS * 16 / B_THREADS;             // staged dwords per thread
    constexpr int BATCH = NE < 32 ? NE : 32;              // loads in flight per thread
    static_assert(NE % BATCH == 0, "whole batches");
    __shared__ uint8_t raw[4 * RQ * G_STRIDE];
    const int tid = threadIdx.x;
    const int64_t row_bytes = nchans * NBITS / 8;
    const int64_t nby = (row_bytes + 63) / 64;
    const int64_t b = blockIdx.x;
    const int64_t logical = (b % 8) * (gridDim.x / 8) + b / 8;
    if (logical >= nby * ((glen + B_TJ - 1) / B_TJ)) return;
    const int64_t j0 = (logical / nby) * B_TJ;
    const int64_t b0 = (logical % nby) * 64;
    // Stage: element e = (row r, dword q of the row's 64 bytes), loaded BATCH at a time.
#pragma unroll 1
    for (int e0 = 0; e0 < NE; e0 += BATCH) {
        uint32_t v[BATCH];
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            const int e = tid + B_THREADS * (e0 + i);
            const int r = e >> 4, q = e & 15;
            const int64_t row = j0 + static_cast<int64_t>(r / B_TJ) * SW_S + (r % B_TJ);
            v[i] = 0;
            if (row < nsamps) {
                const uint8_t* src = in + row * row_bytes + b0 + 4 * q;
                const int64_t left = row_bytes - (b0 + 4 * q);
                if (ALIGNED) {
                    // row_bytes and the stream start are multiples of 4: whole dwords or none
                    if (left > 0) v[i] = *reinterpret_cast<const uint32_t*>(src);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (u < left) v[i] |= static_cast<uint32_t>(src[u]) << (8 * u);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            const int e = tid + B_THREADS * (e0 + i);
            const int r = e >> 4, q = e & 15;
            *reinterpret_cast<uint32_t*>(&raw[((r & 3) * RQ + (r >> 2)) * G_STRIDE + 4 * q]) = v[i];
        }
    }
    __syncthreads();
    const int qd = tid & 31;
    const int64_t j = j0 + 4 * qd;
    if (j >= glen) return;
    for (int bc = tid >> 5; bc < 64; bc += B_THREADS / 32) {
        uint32_t w[PER][4];
#pragma unroll
        for (int p = 0; p < PER; ++p)
#pragma unroll
            for (int u = 0; u < 4; ++u) w[p][u] = 0;
#pragma unroll
        for (int k = 0; k < IL; ++k)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                // row k * B_TJ + 4qd + u
                const uint32_t byte = raw[(u * RQ + k * (B_TJ / 4) + qd) * G_STRIDE + bc];
#pragma unroll
                for (int p = 0; p < PER; ++p) {
                    const int shift = (NBITS == 8) ? 0 : (BIG ? (8 - (p + 1) * NBITS) : p * NBITS);
                    w[p][u] |= ((byte >> shift) & F::MAXV) << (F::FB * k);
                }
            }
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int64_t c = (b0 + bc) * PER + p;
            if (c < nchans) {
                // Streaming store: G (GiBs) is re-read by the sweep much later, not from cache.
                const v4u32 v = {w[p][0], w[p][1], w[p][2], w[p][3]};
                __builtin_nontemporal_store(v, reinterpret_cast<v4u32*>(G + c * gpitch + j));
            }
        }
    }
}

// Stage 2. Per stage of CC channels: nibble data (IL == 8) is summed in NG groups of GN
// channels (GN * maxval <= 15 keeps nibbles exact); each pair of groups is widened to byte
// lanes with one v_add3 per byte plane. Byte data (4-bit) is summed over the stage in byte
// lanes, then widened; 8-bit data goes straight to 16-bit lanes.
template <int NBITS, int CC>
struct StageGroups {
    static constexpr int IL = Fields<NBITS>::IL;
    static constexpr uint32_t MAXV = Fields<NBITS>::MAXV;
    static constexpr int pick() {
        if (IL != 8) return 1;
        for (int ng = 2; ng <= CC; ng += 2)
            if (CC % ng == 0 && static_cast<uint32_t>(CC / ng) * MAXV <= 15u) return ng;
        return -1;
    }
    static constexpr int NG = pick();
    static constexpr int GN = NG > 0 ? CC / NG : 1;
};

// Word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format, no swizzle.
constexpr int kBufferFlags = 0x00020000;

// Bytes 1 and 3 of x into the low bytes of its two 16-bit lanes.
__device__ __forceinline__ uint32_t odd_bytes(uint32_t x) { return __builtin_amdgcn_perm(x, x, 0x0c030c01u); }

template <int NBITS, bool USE32, int K, int WAVES, int CC, int MINW>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(MINW))) void k_sweep(
    const uint32_t* __restrict__ G, int64_t gpitch, int64_t glen, int64_t nchans,
    const uint32_t* __restrict__ relw4, int64_t rpitch4, const int32_t* __restrict__ basew,
    const int32_t* __restrict__ tile_dm0, const int32_t* __restrict__ tile_ndm,
    const int32_t* __restrict__ tile_spread, const int32_t* __restrict__ row_len,
    float* __restrict__ plane, int64_t ld, int64_t shrink, int64_t ntiles_dm, int64_t nblocks) {
    using F = Fields<NBITS>;
    using SG = StageGroups<NBITS, CC>;
    constexpr int IL = F::IL;
    constexpr uint32_t MAXV = F::MAXV;
    constexpr int THREADS = WAVES * 64;
    constexpr int D = K * WAVES;
    constexpr int T = IL * SW_S;
    constexpr int CPW = (CC + WAVES - 1) / WAVES;      // channels each wave stages per stage
    constexpr int NLD = (SW_LW + 255) / 256;            // uint4 loads per lane per staged channel
    constexpr int CHW = SW_LW;                          // words per staged channel window
    constexpr int BUFW = CC * CHW;                      // words per LDS buffer
    static_assert(SW_LW % 4 == 0, "whole uint4 chunks");
    static_assert(D == SW_DT, "the planner forms tiles of SW_DT rows");
    static_assert(CC % 4 == 0 && CC <= 32, "relw rows hold 4 offsets per word, padded for two stages");
    constexpr int NPK = CC / 4;                         // packed offset words per row and stage
    static_assert(IL != 8 || (SG::NG > 0 && SG::NG % 2 == 0), "nibble groups come in pairs");
    static_assert(NBITS != 4 || CC * 15 <= 255, "4-bit stage sums must fit byte lanes");
    constexpr int NG = SG::NG, GN = SG::GN;
    // Stages between byte -> 16-bit widenings (nibble data), and between 16 -> 32-bit ones.
    constexpr int SB = (IL == 8) ? static_cast<int>(255u / (CC * MAXV)) : 1;
    constexpr int SW16 = static_cast<int>(65535u / (CC * MAXV));
    static_assert(SB >= 1 && SW16 >= 1, "lanes must hold one stage");

    __shared__ uint32_t lds[2 * BUFW];

    // XCD-aware order: consecutive blocks on one XCD walk the DM tiles of one time tile.
    const int64_t b = blockIdx.x;
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t logical = (b % 8) * per_xcd + b / 8;
    if (logical >= nblocks) return;
    const int64_t dt = logical % ntiles_dm;
    const int64_t t0 = (logical / ntiles_dm) * T;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int dm0 = tile_dm0[dt];
    const int ndm_t = tile_ndm[dt];
    const int stage_words = SW_S + tile_spread[dt] + 4;
    const int32_t* __restrict__ tbase = basew + dt * nchans;
    const int nst = static_cast<int>((nchans + CC - 1) / CC);

    constexpr int NP = (IL == 8) ? 2 : 1;  // byte planes of nibble data
    uint32_t accB[IL == 8 ? K : 1][4][NP];
    uint32_t accW[K][4][NP][2];
    uint32_t accD[USE32 ? K : 1][USE32 ? 4 * IL : 1];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if constexpr (IL == 8) accB[k][q][p] = 0;
                accW[k][q][p][0] = 0;
                accW[k][q][p][1] = 0;
            }
    if constexpr (USE32) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int i = 0; i < 4 * IL; ++i) accD[k][i] = 0;
    }

    // Nibble data: accB[.][.][1] holds the odd fields (byte lanes); accB[.][.][0] the plain
    // 32-bit sum of the group words, which wraps. As integers the group words sum to
    // even + 16 * odd, and both byte-lane vectors are below 2^32 while their lanes are at most
    // 255, so even = raw - (odd << 4) exactly modulo 2^32.
    auto flushB = [&]() {
        if constexpr (IL == 8) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t odd = accB[k][q][1];
                    const uint32_t even = accB[k][q][0] - (odd << 4);
                    accW[k][q][0][0] += even & 0x00ff00ffu;
                    accW[k][q][0][1] += odd_bytes(even);
                    accW[k][q][1][0] += odd & 0x00ff00ffu;
                    accW[k][q][1][1] += odd_bytes(odd);
                    accB[k][q][0] = 0;
                    accB[k][q][1] = 0;
                }
        }
    };
    // Sample index (0..IL-1) of 16-bit lane l of accW[.][.][p][h].
    auto field_of = [](int p, int h, int l) { return (IL == 8) ? (4 * l + 2 * h + p) : (2 * l + h); };
    // 8-bit data: accW[.][.][0][1] holds the odd bytes (16-bit lanes); accW[.][.][0][0] the plain
    // 32-bit sum of the words, which wraps. A word is even + 256 * odd as integers, and both
    // 16-bit-lane vectors stay below 2^32 while their lanes are below 2^16, so
    // even = raw - (odd << 8) exactly modulo 2^32. evenW() turns the raw sums into the even lanes.
    auto evenW = [&]() {
        if constexpr (NBITS == 8) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q) accW[k][q][0][0] -= accW[k][q][0][1] << 8;
        }
    };
    auto flushW = [&]() {
        if constexpr (USE32) {
            evenW();
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int p = 0; p < NP; ++p)
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            accD[k][q * IL + field_of(p, h, 0)] += accW[k][q][p][h] & 0xffffu;
                            accD[k][q * IL + field_of(p, h, 1)] += accW[k][q][p][h] >> 16;
                            accW[k][q][p][h] = 0;
                        }
        }
    };

    // Staging: wave w copies channels w, w + WAVES, ... of a stage (channel indices, bounds
    // and source rows are wave-uniform); lane l loads words 4l + 256i of each window
    // G[c][t0 + base_c .. + stage_words). Tile bases (basew) are loaded one stage before they
    // are used. The offset of (DM row, channel) inside its window, d - base (t0, a multiple of
    // 4, drops out; at most SW_SPREAD_MAX + 3 < 256), is one byte of relw4 [row][rpitch4 words];
    // each wave reads its own rows' bytes with scalar loads one stage ahead. Rows are
    // zero-padded past nchans (by more than two stages of at most 32 channels), so the last
    // stage's missing channels read their zero-filled windows at offset 0.
    int32_t bw_cur[CPW], bw_nxt[CPW];
    auto load_bw = [&](int st, int32_t* bw) {
#pragma unroll
        for (int i = 0; i < CPW; ++i) {
            const int cl = wave + WAVES * i;
            const int64_t c = static_cast<int64_t>(st) * CC + cl;
            bw[i] = (st < nst && cl < CC && c < nchans) ? tbase[c] : 0;
        }
    };
    // Stage st's windows go straight from G into LDS buffer buf (buffer loads with the LDS
    // flag: lane l's 16 bytes land at the wave-uniform destination + 16 l, the layout the
    // reads below expect), with no register round trip. The buffer was last read in stage
    // st - 1, before the barrier that ended it; the loads are waited for (vmcnt) before the
    // barrier that ends stage st.
    auto prefetch = [&](int st, int buf) {
        const int64_t c0 = static_cast<int64_t>(st) * CC;
#pragma unroll
        for (int i = 0; i < CPW; ++i) {
            const int cl = wave + WAVES * i;
            if (cl >= CC) break;
            const int64_t c = c0 + cl;
            const bool live = c < nchans;
            // Row c of G as a buffer of its glen valid words: reads past them return 0 (a
            // channel past the end has an empty buffer), so no bounds test is needed here.
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<uint32_t*>(G) + (live ? c : 0) * gpitch, 0,
                live ? static_cast<int>(glen * 4) : 0, kBufferFlags);
            const uint32_t off = static_cast<uint32_t>(t0 + bw_cur[i]) * 4u + 16u * lane;
#pragma unroll
            for (int r = 0; r < NLD; ++r) {
                if (4 * lane + 256 * r < stage_words)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rs, (lds_void_t*)(&lds[buf * BUFW + cl * CHW + 256 * r]), 16,
                        off + 1024u * r, 0, 0, 0);
            }
        }
    };
    // vmcnt(0): the stage's LDS-bound loads have landed (expcnt and lgkmcnt left at their
    // maxima, i.e. not waited for).
    auto wait_loads = [&]() { __builtin_amdgcn_s_waitcnt(0x0f70); };

    const char* __restrict__ Lb = reinterpret_cast<const char*>(lds) + 4 * lane;
    const uint32_t* __restrict__ rrow[K];
    uint32_t P[K][NPK];  // this stage's packed offsets of the wave's rows
#pragma unroll
    for (int k = 0; k < K; ++k) {
        rrow[k] = relw4 + static_cast<int64_t>(dm0 + min(wave * K + k, ndm_t - 1)) * rpitch4;
#pragma unroll
        for (int u = 0; u < NPK; ++u) P[k][u] = rrow[k][u];
    }

    load_bw(0, bw_cur);
    load_bw(1, bw_nxt);
    prefetch(0, 0);
    wait_loads();
    __syncthreads();
    int nb = 0, nw = 0;
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) {
#pragma unroll
            for (int i = 0; i < CPW; ++i) bw_cur[i] = bw_nxt[i];
            load_bw(st + 2, bw_nxt);
            prefetch(st + 1, buf ^ 1);
        }
        // The next stage's offsets, waited for at the next barrier (one stage past the last
        // reads the zero padding).
        uint32_t Pn[K][NPK];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < NPK; ++u) Pn[k][u] = rrow[k][static_cast<int64_t>(st + 1) * NPK + u];
        auto rel = [&](int k, int cl) -> uint32_t { return (P[k][cl >> 2] >> (8 * (cl & 3))) & 0xffu; };
        const char* __restrict__ Lbuf = Lb + 4 * buf * BUFW;
        // The 4 words (lane + 64q) of DM row k, stage channel cl.
        auto ld4 = [&](int k, int cl, uint32_t (&w)[4]) {
            const char* p = Lbuf + 4 * cl * CHW + 4 * rel(k, cl);
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = *reinterpret_cast<const uint32_t*>(p + 256 * q);
        };
        // Sum of channels [c0, c0 + n) of DM row k, per word. Loops over channels are fully
        // unrolled (offsets are per-channel scalars); long sums are cut into runs of 4
        // channels the scheduler may not interleave, which bounds the reads in flight.
        auto group_sum = [&](int k, int c0, int n, uint32_t (&s)[4]) {
            ld4(k, c0, s);
#pragma unroll
            for (int j = 1; j < n; ++j) {
                uint32_t w[4];
                ld4(k, c0 + j, w);
#pragma unroll
                for (int q = 0; q < 4; ++q) s[q] += w[q];
                if (n > 5 && j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
            }
        };
        if constexpr (IL == 8) {
#pragma unroll
            for (int g = 0; g < NG; g += 2)
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    uint32_t a[4], bb[4];
                    group_sum(k, g * GN, GN, a);
                    group_sum(k, (g + 1) * GN, GN, bb);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        accB[k][q][0] += a[q] + bb[q];
                        accB[k][q][1] += ((a[q] >> 4) & 0x0f0f0f0fu) + ((bb[q] >> 4) & 0x0f0f0f0fu);
                    }
                    // Keep the scheduler from hoisting every row's LDS reads (register spills).
                    __builtin_amdgcn_sched_barrier(0);
                }
        } else {
            // Byte fields: a runtime loop over the stage's packed offset words, 4 channels per
            // trip (short live ranges). The words are consumed from a shift register so every
            // index stays static.
#pragma unroll
            for (int k = 0; k < K; ++k) {
                uint32_t pk[NPK];
#pragma unroll
                for (int u = 0; u < NPK; ++u) pk[u] = P[k][u];
                uint32_t a[4] = {0, 0, 0, 0};
#pragma unroll 1
                for (int u = 0; u < NPK; ++u) {
                    const uint32_t word = pk[0];
#pragma unroll
                    for (int v = 0; v + 1 < NPK; ++v) pk[v] = pk[v + 1];
                    const char* base = Lbuf + 16 * u * CHW;
                    uint32_t w[4][4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const char* pj = base + 4 * j * CHW + 4 * ((word >> (8 * j)) & 0xffu);
#pragma unroll
                        for (int q = 0; q < 4; ++q) w[j][q] = *reinterpret_cast<const uint32_t*>(pj + 256 * q);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if constexpr (NBITS == 4) {
                            a[q] += w[0][q] + w[1][q] + w[2][q] + w[3][q];  // byte lanes, <= CC * 15
                        } else {
                            accW[k][q][0][0] += w[0][q] + w[1][q];  // raw sum (evenW)
                            accW[k][q][0][0] += w[2][q] + w[3][q];
                            accW[k][q][0][1] += odd_bytes(w[0][q]) + odd_bytes(w[1][q]);
                            accW[k][q][0][1] += odd_bytes(w[2][q]) + odd_bytes(w[3][q]);
                        }
                    }
                }
                if constexpr (NBITS == 4) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        accW[k][q][0][0] += a[q] & 0x00ff00ffu;
                        accW[k][q][0][1] += odd_bytes(a[q]);
                    }
                }
            }
        }
        if constexpr (IL == 8) {
            if (++nb == SB) {
                flushB();
                nb = 0;
            }
        }
        if constexpr (USE32) {
            if (++nw == SW16) {
                flushB();
                nb = 0;
                flushW();
                nw = 0;
            }
        }
        if (st + 1 < nst) wait_loads();
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < NPK; ++u) P[k][u] = Pn[k][u];
        __syncthreads();
    }
    flushB();
    evenW();

    // Store: field f of word (lane + 64q) is sample t0 + lane + 64q + f*S of DM row dm0 + wave*K + k.
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int idx = wave * K + k;
        if (idx >= ndm_t) break;
        const int64_t row = dm0 + idx;
        const int64_t len = row_len[row] - shrink;
        float* __restrict__ orow = plane + row * ld;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int l = 0; l < 2; ++l) {
                        const int f = field_of(p, h, l);
                        const int64_t t = t0 + lane + 64 * q + static_cast<int64_t>(SW_S) * f;
                        uint32_t v = (l == 0) ? (accW[k][q][p][h] & 0xffffu) : (accW[k][q][p][h] >> 16);
                        if constexpr (USE32) v += accD[k][q * IL + f];
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
    std::vector<int32_t> base;   // [ntiles][nchans] = min delay of the tile, rounded down to 4
    std::vector<int32_t> row_len;
    int64_t max_len = 0, min_len = 0, total = 0;
    int32_t max_spread = 0;
    bool use32 = false;
    int variant = 0;  // k_sweep block shape (launch_sweep)
    int64_t gpitch = 0;
    std::vector<uint8_t> relw;   // [ndm][rpitch] = delay - tile base (rounded down to 4) < 256
    int64_t rpitch = 0;          // bytes per row: nchans + 64 rounded up to 16; zero past nchans
    int64_t off_relw = 0, off_base = 0, off_dm0 = 0, off_ndm = 0, off_spread = 0, off_rowlen = 0,
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
    // 8 waves x 2 DM rows x 12 channels per stage at 6 waves per SIMD: 48 KB of LDS, three
    // blocks per CU (nibble data: variant 3, byte data: variant 0); tunable per plan
    p->variant = nbits <= 2 ? 3 : 0;
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

extern "C" int spp_sweep_variant_count(void) { return 4; }

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
    return SPP_OK;
}

namespace {

// One execution of a plan on a stream of n <= plan nsamps samples: every row is shorter by
// shrink = nsamps - n, and G holds glen valid words per channel.
struct Extent {
    int64_t n, glen, shrink, max_len;
};

Extent extent_of(const spp_sweep_plan* p, int64_t n) {
    const int64_t glen = std::min(p->gpitch, ((n + 63) / 64) * 64);
    return Extent{n, glen, p->nsamps - n, p->max_len - (p->nsamps - n)};
}

template <int NBITS, bool BIG>
int launch_build(const uint8_t* packed, const spp_sweep_plan* p, const Extent& e, uint32_t* G,
                 hipStream_t st) {
    const int64_t row_bytes = p->nchans * NBITS / 8;
    const int64_t nblocks = spp::ceil_div(e.glen, int64_t(B_TJ)) * spp::ceil_div(row_bytes, int64_t(64));
    SPP_REQUIRE(nblocks + 8 <= 0x7fffffffLL, "stage-1 grid too large");
    const dim3 grid(static_cast<unsigned>(spp::ceil_div(nblocks, int64_t(8)) * 8));
    if (row_bytes % 4 == 0 && reinterpret_cast<uintptr_t>(packed) % 4 == 0)
        hipLaunchKernelGGL((k_build_g<NBITS, BIG, true>), grid, dim3(B_THREADS), 0, st, packed, e.n, p->nchans, G,
                           p->gpitch, e.glen);
    else
        hipLaunchKernelGGL((k_build_g<NBITS, BIG, false>), grid, dim3(B_THREADS), 0, st, packed, e.n, p->nchans, G,
                           p->gpitch, e.glen);
    SPP_LAUNCHED("k_build_g");
    return SPP_OK;
}

template <int NBITS, bool USE32, int K, int WAVES, int CC, int MINW>
int launch_sweep_v(const spp_sweep_plan* p, const Extent& e, char* ws, float* plane, int64_t ld,
                   hipStream_t st) {
    const int64_t ntiles_dm = static_cast<int64_t>(p->tile_dm0.size());
    const int64_t ntiles_t = spp::ceil_div(e.max_len, int64_t(Fields<NBITS>::IL) * SW_S);
    const int64_t nblocks = ntiles_dm * ntiles_t;
    const int64_t grid = spp::ceil_div(nblocks, 8) * 8;
    SPP_REQUIRE(grid <= 0x7fffffffLL, "sweep grid too large");
    hipLaunchKernelGGL((k_sweep<NBITS, USE32, K, WAVES, CC, MINW>), dim3(static_cast<unsigned>(grid)),
                       dim3(WAVES * 64), 0, st, reinterpret_cast<const uint32_t*>(ws + p->off_g),
                       p->gpitch, e.glen, p->nchans, reinterpret_cast<const uint32_t*>(ws + p->off_relw),
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
// Nibble data (1/2-bit): 20 channels per stage give 5-channel nibble groups for 2-bit data.
// 6 waves per SIMD (<= 80 VGPRs: two accumulators live in scratch, touched only at the byte ->
// 16-bit widenings, the rest of the spills sit in the epilogue) with 12 channels per stage
// (48 KB of LDS) fits three blocks per CU and is the fastest shape measured.
template <int NBITS, bool USE32>
int launch_sweep(const spp_sweep_plan* p, const Extent& e, char* ws, float* plane, int64_t ld,
                 hipStream_t st) {
    if constexpr (NBITS <= 2) {
        switch (p->variant) {
            case 1: return launch_sweep_v<NBITS, USE32, 2, 8, 20, 4>(p, e, ws, plane, ld, st);
            case 2: return launch_sweep_v<NBITS, USE32, 2, 8, 8, 6>(p, e, ws, plane, ld, st);
            case 3: return launch_sweep_v<NBITS, USE32, 2, 8, 12, 6>(p, e, ws, plane, ld, st);
            default: return launch_sweep_v<NBITS, USE32, 2, 8, 16, 4>(p, e, ws, plane, ld, st);
        }
    } else {
        switch (p->variant) {
            case 1: return launch_sweep_v<NBITS, USE32, 2, 8, 16, 4>(p, e, ws, plane, ld, st);
            case 2: return launch_sweep_v<NBITS, USE32, 4, 4, 8, 3>(p, e, ws, plane, ld, st);
            case 3: return launch_sweep_v<NBITS, USE32, 2, 8, 8, 4>(p, e, ws, plane, ld, st);
            default: return launch_sweep_v<NBITS, USE32, 2, 8, 12, 6>(p, e, ws, plane, ld, st);
        }
    }
}

}  // namespace

namespace {

int check_exec(const spp_sweep_plan* p, const void* workspace, int64_t workspace_bytes) {
    SPP_REQUIRE(p != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace == p->uploaded_to, "workspace is not the buffer the plan was uploaded to");
    SPP_REQUIRE(workspace_bytes >= p->ws_bytes, "workspace too small");
    return SPP_OK;
}

int build_n(const spp_sweep_plan* p, const uint8_t* packed, const Extent& e, void* workspace,
            hipStream_t st) {
    uint32_t* G = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + p->off_g);
    switch (p->nbits * 2 + p->big) {
        case 2: return launch_build<1, false>(packed, p, e, G, st);
        case 3: return launch_build<1, true>(packed, p, e, G, st);
        case 4: return launch_build<2, false>(packed, p, e, G, st);
        case 5: return launch_build<2, true>(packed, p, e, G, st);
        case 8: return launch_build<4, false>(packed, p, e, G, st);
        case 9: return launch_build<4, true>(packed, p, e, G, st);
        default: return launch_build<8, false>(packed, p, e, G, st);
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
    return build_n(p, packed, extent_of(p, p->nsamps), workspace, spp::as_stream(stream));
}

extern "C" int spp_sweep_run(const spp_sweep_plan* p, float* plane, int64_t ld, void* workspace,
                             int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_exec(p, workspace, workspace_bytes);
    if (rc) return rc;
    return run_n(p, extent_of(p, p->nsamps), plane, ld, workspace, spp::as_stream(stream));
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
    const Extent e = extent_of(p, nsamps);
    hipStream_t st = spp::as_stream(stream);
    rc = build_n(p, packed, e, workspace, st);
    if (rc) return rc;
    return run_n(p, e, plane, ld, workspace, st);
}
