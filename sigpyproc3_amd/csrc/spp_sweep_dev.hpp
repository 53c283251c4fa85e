// Device code of the DM-trial sweep, shared by the one-plan kernels (sweep.hip) and the
// persistent batch kernel (batch.hip).
//
// Stage 1 (build_tile) turns the packed time-major stream [nsamps][nchans] into a
// channel-major, time-interleaved word array G[c][j] (uint32) holding IL samples
// x_c[j + k*S] (k < IL, S = 256) in fields of 32/IL bits, zero past the end:
//   1- and 2-bit data: IL = 8 nibble fields; 4- and 8-bit data: IL = 4 byte fields.
// A delay shift d of channel c moves every field of a word together, so an output word
// (the IL samples t + k*S) is the sum over channels of the aligned words G[c][t + d_i(c)].
//
// Stage 2 (spp_sweep_tile.inc): a block owns D <= 16 rows of a DM tile (<= 16 consecutive DM rows
// whose per-channel delay spread is <= SW_SPREAD_MAX, formed by the host planner) and IL*256
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
#pragma once

#include "spp_common.hpp"

namespace {

typedef __attribute__((address_space(3))) void lds_void_t;  // an LDS address for the LDS-load builtin

constexpr int SW_S = 256;                // interleave stride = words per wave row
constexpr int SW_DT = 16;                // DM rows per tile
constexpr int SW_LW = 512;               // LDS words per staged channel (2 KiB: a multiple of the
                                         // 256-byte unit of ds_read2st64 offsets)
constexpr int SW_SPREAD_MAX = SW_LW - SW_S - 4;  // per-channel delay spread of a DM tile
constexpr int B_TJ = 128;                // words per build tile (time)
constexpr int B_THREADS = 512;           // threads per build tile

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

template <int NBITS>
struct Fields {
    static constexpr int IL = NBITS <= 2 ? 8 : 4;          // samples per word
    static constexpr int FB = 32 / IL;                     // bits per field
    static constexpr uint32_t MAXV = NBITS == 8 ? 255u : ((1u << NBITS) - 1u);
};

// LDS bytes build_tile<NBITS, ., ., BCW> stages.
template <int NBITS, int BCW>
constexpr int build_lds_bytes() {
    return 4 * (Fields<NBITS>::IL * B_TJ / 4 + 1) * (BCW + 4);
}

// Stage 1, one tile: words [j0, j0 + B_TJ) of the channels in row bytes [b0, b0 + BCW)
// (BCW * 8/NBITS channels), by B_THREADS threads. The rows j0 + k*S + i (k < IL, i < B_TJ)
// are staged in LDS (`raw`, build_lds_bytes() bytes). Thread (quad q, byte column bc) then
// builds words j0 + 4q .. 4q+3 of every channel packed in byte column bc from one LDS byte
// per (row, word) and stores them as one 16-byte write per channel (a wave writes 512
// contiguous bytes per channel). Only the first glen words of each channel row are written
// (the sweep reads no further). One barrier, reached by every thread of the block. MAXB bounds
// the staging loads in flight per thread (registers); tid is the thread's index in the block
// of NT threads (B_THREADS or a multiple).
template <int NBITS, bool BIG, bool ALIGNED, int BCW, int MAXB = 32, int NT = B_THREADS>
__device__ __forceinline__ void build_tile(const uint8_t* __restrict__ in, int64_t nsamps, int64_t nchans,
                                           uint32_t* __restrict__ G, int64_t gpitch, int64_t glen,
                                           int64_t j0, int64_t b0, uint8_t* raw, int tid) {
    using F = Fields<NBITS>;
    constexpr int IL = F::IL;
    constexpr int PER = (NBITS < 8) ? 8 / NBITS : 1;     // channels per byte
    constexpr int ROWS = IL * B_TJ;
    constexpr int QPR = BCW / 4;                          // staged dwords per row
    // Staged row r sits at position (r % 4) * RQ + r / 4, STRIDE bytes per row (an odd number
    // of dwords): the rows 4qd + u a build step reads for one u are then consecutive and fall
    // on distinct banks.
    constexpr int RQ = ROWS / 4 + 1;
    constexpr int STRIDE = BCW + 4;
    constexpr int NE = ROWS * QPR / NT;                   // staged dwords per thread
    constexpr int BATCH = NE < MAXB ? NE : MAXB;          // loads in flight per thread
    static_assert(BCW % 16 == 0 && BCW <= 64, "16, 32, 48 or 64 byte columns");
    static_assert(NE >= 1 && NE % BATCH == 0 && NE * NT == ROWS * QPR, "whole batches");
    const int64_t row_bytes = nchans * NBITS / 8;
    // Stage: element e = (row r, dword q of the row's BCW bytes), loaded BATCH at a time.
#pragma unroll 1
    for (int e0 = 0; e0 < NE; e0 += BATCH) {
        uint32_t v[BATCH];
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
            const int e = tid + NT * (e0 + i);
            const int r = e / QPR, q = e % QPR;
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
            const int e = tid + NT * (e0 + i);
            const int r = e / QPR, q = e % QPR;
            *reinterpret_cast<uint32_t*>(&raw[((r & 3) * RQ + (r >> 2)) * STRIDE + 4 * q]) = v[i];
        }
    }
    __syncthreads();
    const int qd = tid & 31;
    const int64_t j = j0 + 4 * qd;
    if (j >= glen) return;
    for (int bc = tid >> 5; bc < BCW; bc += NT / 32) {
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
                const uint32_t byte = raw[(u * RQ + k * (B_TJ / 4) + qd) * STRIDE + bc];
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

}  // namespace
