// Gulp pre-steps that share the dedispersion path's reads and unpack: per-sample channel sums
// (collapse), per-channel sums (bandpass), channel masking on the file's own bytes, and
// zero-DM removal.
//
// Gulp layout: file order [nsamps][nchans], packed nbits (1/2/4/8, bit order `big`) or
// float32 (nbits == 32). Each is a single pass over the gulp (bandpass: plus a pass over its
// per-chunk partial sums).
#include "spp_common.hpp"
#include "spp_unpack.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace {

template <int NBITS, bool BIG>
__device__ __forceinline__ float sample_at(const void* in, int64_t row, int64_t row_bytes, int64_t c) {
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

// ---------------------------------------------------------------------------------------
// Integer gulps (nbits 1/2/4/8). Every sample is a small non-negative integer, so a float32
// running sum that starts at an integer o with |o| <= 2^24 and ends at |o + T| <= 2^24 never
// rounds: every partial sum is an integer of at most 24 bits. In that range any summation
// order gives the reference's float32 result, and the sums below are exact integer
// reductions in whatever order the hardware prefers.

// Sum of every field of a packed word (bit order does not matter for a sum): 1-bit popcount,
// 2-bit popcount of the word plus popcount of the high bits, 4/8-bit byte SAD against zero.
template <int NBITS>
__device__ __forceinline__ uint32_t field_sum(uint32_t w, uint32_t acc) {
    if constexpr (NBITS == 1) {
        return acc + static_cast<uint32_t>(__builtin_popcount(w));
    } else if constexpr (NBITS == 2) {
        return acc + static_cast<uint32_t>(__builtin_popcount(w)) +
               static_cast<uint32_t>(__builtin_popcount(w & 0xaaaaaaaau));
    } else if constexpr (NBITS == 4) {
        return __builtin_amdgcn_sad_u8(w & 0x0f0f0f0fu, 0u,
                                       __builtin_amdgcn_sad_u8((w >> 4) & 0x0f0f0f0fu, 0u, acc));
    } else {
        return __builtin_amdgcn_sad_u8(w, 0u, acc);
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o));
    return v;
}

// Collapse (kernels.extract_tim, core/kernels.py:255-273) of an integer gulp: one wave per
// CL_ROWS rows, UB bytes per lane load; out[t] = float32 of the exact row sum (the
// reference's float64 sum of integers, rounded once on the store).
constexpr int CL_ROWS = 4;
template <int NBITS, int UB>
__global__ __launch_bounds__(256) void k_collapse(const uint8_t* __restrict__ in, int64_t nsamps,
                                                  int64_t row_bytes, float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t r0 = (static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * CL_ROWS;
    const int64_t nunit = row_bytes / UB;
    uint32_t s[CL_ROWS] = {};
    for (int64_t u = lane; u < nunit; u += 64) {
        uint32_t v[CL_ROWS][UB >= 4 ? UB / 4 : 1];
#pragma unroll
        for (int i = 0; i < CL_ROWS; ++i) {
            const int64_t r = min(r0 + i, nsamps - 1);  // rows past the end re-read the last one
            const uint8_t* p = in + r * row_bytes + u * UB;
            if constexpr (UB == 16) {
                const uint4 q = *reinterpret_cast<const uint4*>(p);
                v[i][0] = q.x;
                v[i][1] = q.y;
                v[i][2] = q.z;
                v[i][3] = q.w;
            } else if constexpr (UB == 4) {
                v[i][0] = *reinterpret_cast<const uint32_t*>(p);
            } else {
                v[i][0] = *p;
            }
        }
#pragma unroll
        for (int i = 0; i < CL_ROWS; ++i) {
#pragma unroll
            for (int k = 0; k < (UB >= 4 ? UB / 4 : 1); ++k) s[i] = field_sum<NBITS>(v[i][k], s[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < CL_ROWS; ++i) {
        const uint32_t t = wave_sum(s[i]);
        if (lane == 0 && r0 + i < nsamps) out[r0 + i] = static_cast<float>(t);
    }
}

// Collapse of float32 rows: one thread per row, float64 sum in channel order.
__global__ __launch_bounds__(256) void k_collapse_f32(const float* __restrict__ in, int64_t nsamps,
                                                      int64_t nchans, float* __restrict__ out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= nsamps) return;
    const float* row = in + t * nchans;
    double s = 0.0;
    for (int64_t c = 0; c < nchans; ++c) s += static_cast<double>(row[c]);
    out[t] = static_cast<float>(s);
}

// Bandpass, pass 1: per-channel integer sums of one chunk of rows. A block covers 64 lane
// units of a row (VW dwords each, or one byte when VW == 0) and chunk_rows rows; its four
// waves take every fourth row. Samples are added as SWAR fields (8-bit lanes for 1/2-bit
// data, 16-bit lanes for 4/8-bit), widened into per-channel counters before a lane can
// overflow, and the waves' counters are combined in LDS. part[chunk][c] is the chunk's sum.
constexpr int BP_R = 8;          // rows in flight per lane
constexpr int BP_BLOCKS = 1024;  // target pass-1 grid
template <int NBITS, bool BIG, int VW>
__global__ __launch_bounds__(256) void k_bp_part(const uint8_t* __restrict__ in, int64_t nsamps,
                                                 int64_t row_bytes, int64_t chunk_rows, int64_t nchans,
                                                 uint32_t* __restrict__ part) {
    constexpr int PER = 8 / NBITS;
    constexpr int NW = VW ? VW : 1;          // words per lane unit
    constexpr int UB = VW ? 4 * VW : 1;      // bytes per lane unit
    constexpr int NB = VW ? 4 : 1;           // bytes used per word
    constexpr int CW = NB * PER;             // channels per word
    constexpr int CPU = NW * CW;             // channels per lane unit
    constexpr bool BL = NBITS <= 2;          // 8-bit lanes (else 16-bit lanes)
    constexpr int NA = BL ? PER : 2 * PER;   // SWAR accumulators per word
    constexpr int FL = BL ? 64 : 256;        // rows between widenings: 3*64 < 2^8, 255*256 < 2^16
    constexpr uint32_t FM = NBITS == 8 ? 0xffu : (1u << NBITS) - 1u;
    constexpr uint32_t LM = BL ? FM * 0x01010101u : FM * 0x00010001u;
    constexpr int LSTR = CPU + 1;            // odd LDS stride per lane unit
    __shared__ uint32_t red[64 * LSTR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nunit = row_bytes / UB;
    const int64_t unit = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    const bool active = unit < nunit;
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * chunk_rows;
    const int64_t r1 = min(nsamps, r0 + chunk_rows);
    uint32_t acc[NW][NA];
    uint32_t tot[NW][CW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[w][a] = 0;
#pragma unroll
        for (int j = 0; j < CW; ++j) tot[w][j] = 0;
    }
    auto widen = [&]() {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int k = BL ? a : (a >> 1);
                const int ck = BIG ? (PER - 1 - k) : k;
                if constexpr (BL) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) tot[w][b * PER + ck] += (acc[w][a] >> (8 * b)) & 0xffu;
                } else {
#pragma unroll
                    for (int hf = 0; hf < 2; ++hf) {
                        const int b = (a & 1) + 2 * hf;
                        if (b < NB) tot[w][b * PER + ck] += (acc[w][a] >> (16 * hf)) & 0xffffu;
                    }
                }
                acc[w][a] = 0;
            }
        }
    };
    int pend = 0;
    for (int64_t r = r0 + wv; r < r1; r += 4 * BP_R) {
        uint32_t x[BP_R][NW];
#pragma unroll
        for (int i = 0; i < BP_R; ++i) {
            const int64_t row = r + 4 * i;
            const bool ok = active && row < r1;
            const uint8_t* p = in + (ok ? row * row_bytes + unit * UB : 0);
            if constexpr (VW == 4) {
                uint4 q = make_uint4(0, 0, 0, 0);
                if (ok) q = *reinterpret_cast<const uint4*>(p);
                x[i][0] = q.x;
                x[i][1] = q.y;
                x[i][2] = q.z;
                x[i][3] = q.w;
            } else if constexpr (VW == 2) {
                uint2 q = make_uint2(0, 0);
                if (ok) q = *reinterpret_cast<const uint2*>(p);
                x[i][0] = q.x;
                x[i][1] = q.y;
            } else if constexpr (VW == 1) {
                x[i][0] = ok ? *reinterpret_cast<const uint32_t*>(p) : 0u;
            } else {
                x[i][0] = ok ? static_cast<uint32_t>(*p) : 0u;
            }
        }
#pragma unroll
        for (int i = 0; i < BP_R; ++i) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int sh = BL ? a * NBITS : (8 * (a & 1) + (a >> 1) * NBITS);
                    acc[w][a] += (x[i][w] >> sh) & LM;
                }
            }
        }
        pend += BP_R;
        if (pend + BP_R > FL) {
            widen();
            pend = 0;
        }
    }
    widen();
    if (wv == 0) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
#pragma unroll
            for (int j = 0; j < CW; ++j) red[lane * LSTR + w * CW + j] = tot[w][j];
        }
    }
    __syncthreads();
    if (wv != 0) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
#pragma unroll
            for (int j = 0; j < CW; ++j) atomicAdd(&red[lane * LSTR + w * CW + j], tot[w][j]);
        }
    }
    __syncthreads();
    const int64_t cbase = static_cast<int64_t>(blockIdx.x) * 64 * CPU;
    uint32_t* dst = part + static_cast<int64_t>(blockIdx.y) * nchans;
    for (int l = threadIdx.x; l < 64 * CPU; l += 256) {
        const int64_t c = cbase + l;
        if (c < nchans) dst[c] = red[(l / CPU) * LSTR + l % CPU];
    }
}

// Bandpass, pass 2: one channel per lane, 16 waves over the chunks. T = the gulp's exact sum.
//   parallel (kernels.extract_bpass_par): out = float32(float64(out) + T), numba's
//     float32 + float64 -> float64 store;
//   sequential (kernels.extract_bpass): out + T when every running float32 sum stays an
//     integer within 2^24 (then it is exact); otherwise the exact prefix up to the first chunk
//     that could leave that range, then float32 adds in time order from there.
constexpr int BPF_WAVES = 16;
template <int NBITS, bool BIG, bool SEQ>
__global__ __launch_bounds__(64 * BPF_WAVES) void k_bp_finish(const uint32_t* __restrict__ part,
                                                              int64_t nchunks, int64_t chunk_rows,
                                                              int64_t nchans, const uint8_t* __restrict__ in,
                                                              int64_t nsamps, int64_t row_bytes,
                                                              float* __restrict__ out) {
    __shared__ uint64_t red[BPF_WAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    uint64_t s = 0;
    if (c < nchans) {
#pragma unroll 8
        for (int64_t j = wv; j < nchunks; j += BPF_WAVES) s += part[j * nchans + c];
    }
    red[wv][lane] = s;
    __syncthreads();
    if (wv != 0 || c >= nchans) return;
    uint64_t tsum = 0;
#pragma unroll
    for (int k = 0; k < BPF_WAVES; ++k) tsum += red[k][lane];
    const float o = out[c];
    const double total = static_cast<double>(tsum);
    if constexpr (!SEQ) {
        out[c] = static_cast<float>(static_cast<double>(o) + total);
    } else {
        constexpr double LIM = 16777216.0;
        const bool o_int = o == truncf(o) && fabs(static_cast<double>(o)) <= LIM;
        if (o_int && fabs(static_cast<double>(o) + total) <= LIM) {
            out[c] = static_cast<float>(static_cast<double>(o) + total);
            return;
        }
        double pre = o;
        int64_t j = 0;
        if (o_int) {
            while (j < nchunks) {
                const double nx = pre + static_cast<double>(part[j * nchans + c]);
                if (fabs(nx) > LIM) break;
                pre = nx;
                ++j;
            }
        }
        float acc = static_cast<float>(pre);
        int64_t t = j * chunk_rows;
        for (; t + 8 <= nsamps; t += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = sample_at<NBITS, BIG>(in, t + u, row_bytes, c);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + v[u];
        }
        for (; t < nsamps; ++t) acc = acc + sample_at<NBITS, BIG>(in, t, row_bytes, c);
        out[c] = acc;
    }
}

// Bandpass of float32 rows, one thread per channel, time in order.
//   SEQ:  out[c] = (...((out[c] + x[0][c]) + x[1][c]) ...) in float32 (kernels.extract_bpass);
//   !SEQ: out[c] = float32(float64(out[c]) + sum_t x[t][c] in float64, time order)
//         (kernels.extract_bpass_par).
template <bool SEQ>
__global__ __launch_bounds__(256) void k_bandpass_f32(const float* __restrict__ in, int64_t nchans,
                                                      int64_t nsamps, float* __restrict__ out) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (c >= nchans) return;
    using acc_t = std::conditional_t<SEQ, float, double>;
    acc_t acc = SEQ ? static_cast<acc_t>(out[c]) : acc_t(0);
    int64_t t = 0;
    for (; t + 16 <= nsamps; t += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = in[(t + u) * nchans + c];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = acc + static_cast<acc_t>(v[u]);
    }
    for (; t < nsamps; ++t) acc = acc + static_cast<acc_t>(in[t * nchans + c]);
    if constexpr (SEQ) out[c] = acc;
    else out[c] = static_cast<float>(static_cast<double>(out[c]) + acc);
}

// Pass-1 geometry shared by the workspace query and the launch: lane-unit width, grid
// columns, chunk count and chunk length.
struct BpGeom {
    int vw;
    int64_t gx, nchunks, chunk_rows;
};

BpGeom bp_geom(int nbits, int64_t row_bytes, int64_t nsamps, bool aligned16, bool aligned8,
               bool aligned4) {
    BpGeom g{};
    const int vmax = nbits == 1 ? 2 : 4;
    if (vmax == 4 && aligned16 && row_bytes % 16 == 0) g.vw = 4;
    else if (vmax == 2 && aligned8 && row_bytes % 8 == 0) g.vw = 2;
    else if (aligned4 && row_bytes % 4 == 0) g.vw = 1;
    else g.vw = 0;
    const int64_t ub = g.vw ? 4 * g.vw : 1;
    g.gx = spp::ceil_div(row_bytes / ub, int64_t(64));
    // The chunk count comes from the widest unit, so the workspace query (which does not
    // know the alignment) and every launch agree on it.
    const int64_t gx_min = spp::ceil_div(spp::ceil_div(row_bytes, int64_t(4 * vmax)), int64_t(64));
    const int64_t want = std::max<int64_t>(1, spp::ceil_div(BP_BLOCKS, gx_min));
    const int64_t maxc = std::max<int64_t>(1, spp::ceil_div(nsamps, int64_t(4 * BP_R)));
    const int64_t nc = std::min(want, maxc);
    g.chunk_rows = spp::ceil_div(spp::ceil_div(nsamps, nc), int64_t(4 * BP_R)) * (4 * BP_R);
    g.nchunks = spp::ceil_div(nsamps, g.chunk_rows);
    return g;
}

// Masking on the file's bytes: byte j of every row becomes (b & keep[j]) | set[j]. The host
// builds keep/set from the mask and value so this equals kernels.mask_channels
// (core/kernels.py:321-345) on the unpacked samples followed by the reference's pack.
__global__ __launch_bounds__(256) void k_mask_bytes(uint8_t* __restrict__ data, int64_t nbytes,
                                                    int64_t row_bytes, const uint8_t* __restrict__ keep,
                                                    const uint8_t* __restrict__ set) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < nbytes; i += stride) {
        const int64_t j = i % row_bytes;
        data[i] = static_cast<uint8_t>((data[i] & keep[j]) | set[j]);
    }
}

// Zero-DM removal (kernels.remove_zerodm, core/kernels.py:469-500): one block per sample row.
//   z = sum_c x[t][c] (exact for integer data; float64 for float32 data),
//   y = fma(-z, w[c], x[t][c]) + b[c] in float64,
//   out = float32(y) for float32 data, else the low 8 bits of y truncated toward zero (the
//   float64 -> uint8 store of the reference on x86-64).
template <int NBITS, bool BIG>
__global__ __launch_bounds__(256) void k_zerodm(const void* __restrict__ in, int64_t nchans,
                                                int64_t row_bytes, const float* __restrict__ w,
                                                const float* __restrict__ bpass, void* __restrict__ out) {
    __shared__ double part[256];
    const int64_t t = blockIdx.x;
    double s = 0.0;
    for (int64_t c = threadIdx.x; c < nchans; c += 256) s += static_cast<double>(sample_at<NBITS, BIG>(in, t, row_bytes, c));
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    const double z = part[0];
    for (int64_t c = threadIdx.x; c < nchans; c += 256) {
        const double x = static_cast<double>(sample_at<NBITS, BIG>(in, t, row_bytes, c));
        const double y = fma(-z, static_cast<double>(w[c]), x) + static_cast<double>(bpass[c]);
        if constexpr (NBITS == 32) {
            static_cast<float*>(out)[t * nchans + c] = static_cast<float>(y);
        } else {
            // In-range guard keeps the conversion defined; |y| < 2^31 for any real bandpass.
            const double yc = fmin(fmax(y, -2147483648.0), 2147483647.0);
            static_cast<uint8_t*>(out)[t * nchans + c] = static_cast<uint8_t>(static_cast<int32_t>(yc) & 0xff);
        }
    }
}

// Masking with dword or 16-byte units: a thread keeps one unit's keep/set bytes and applies
// them to that unit of rows blockIdx.y, blockIdx.y + gridDim.y, ...
template <int UB>
__global__ __launch_bounds__(256) void k_mask_rows(uint8_t* __restrict__ data, int64_t nrows, int64_t row_bytes,
                                                   const uint8_t* __restrict__ keep,
                                                   const uint8_t* __restrict__ set) {
    using unit_t = std::conditional_t<UB == 16, uint4, uint32_t>;
    const int64_t nunit = row_bytes / UB;
    const int64_t u = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= nunit) return;
    const unit_t k = reinterpret_cast<const unit_t*>(keep)[u];
    const unit_t st = reinterpret_cast<const unit_t*>(set)[u];
    constexpr int R = 4;
    for (int64_t r0 = blockIdx.y; r0 < nrows; r0 += static_cast<int64_t>(gridDim.y) * R) {
        unit_t v[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int64_t r = r0 + static_cast<int64_t>(i) * gridDim.y;
            if (r < nrows) v[i] = reinterpret_cast<const unit_t*>(data + r * row_bytes)[u];
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int64_t r = r0 + static_cast<int64_t>(i) * gridDim.y;
            if (r >= nrows) break;
            unit_t o;
            if constexpr (UB == 16) {
                o = make_uint4((v[i].x & k.x) | st.x, (v[i].y & k.y) | st.y, (v[i].z & k.z) | st.z,
                               (v[i].w & k.w) | st.w);
            } else {
                o = (v[i] & k) | st;
            }
            reinterpret_cast<unit_t*>(data + r * row_bytes)[u] = o;
        }
    }
}

// Zero-DM removal of integer data with dword units (row bytes and pointers 4-byte aligned):
// a thread owns UPT dwords of every row (4 * 8/NBITS channels each, at most 32 channels) and
// keeps their weights and bandpass values in registers; the block walks rows. Per row: z from the
// exact field sums (block reduction), then every sample as in k_zerodm, four output bytes per
// dword store.
template <int NBITS, bool BIG, int UPT>
__global__ __launch_bounds__(256) void k_zerodm_w(const uint8_t* __restrict__ in, int64_t nsamps,
                                                  int64_t nchans, int64_t row_bytes,
                                                  const float* __restrict__ w, const float* __restrict__ bpass,
                                                  uint8_t* __restrict__ out, int vec) {
    constexpr int PER = 8 / NBITS;
    constexpr int CPW = 4 * PER;  // channels per dword
    __shared__ uint32_t zpart[2][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t nunit = row_bytes / 4;
    static_assert(UPT * CPW <= 32, "channels per thread");
    float wf[UPT][CPW], bf[UPT][CPW];
    double wmax = 0.0, bmax = 0.0;
#pragma unroll
    for (int i = 0; i < UPT; ++i) {
        const int64_t u = tid + 256 * i;
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            const int64_t c = u * CPW + j;
            wf[i][j] = u < nunit ? w[c] : 0.f;
            bf[i][j] = u < nunit ? bpass[c] : 0.f;
            wmax = fmax(wmax, fabs(static_cast<double>(wf[i][j])));
            bmax = fmax(bmax, fabs(static_cast<double>(bf[i][j])));
        }
    }
    // |y| <= 255 + z * |w| + |b| with z <= (2^nbits - 1) * nchans: when that stays below 2^30
    // for all of this thread's channels, y converts to int32 without the clamp. (A NaN or
    // infinite weight fails the test and keeps the clamped form.)
    const double zmax = static_cast<double>((NBITS == 8 ? 255 : (1 << NBITS) - 1)) * static_cast<double>(nchans);
    const bool direct = 255.0 + zmax * wmax + bmax < 1073741824.0;
    auto load_row = [&](int64_t t, uint32_t (&v)[UPT]) {
#pragma unroll
        for (int i = 0; i < UPT; ++i) {
            const int64_t u = tid + 256 * i;
            v[i] = (t < nsamps && u < nunit) ? reinterpret_cast<const uint32_t*>(in + t * row_bytes)[u] : 0u;
        }
    };
    int par = 0;
    uint32_t v[UPT];
    load_row(blockIdx.x, v);
    for (int64_t t = blockIdx.x; t < nsamps; t += gridDim.x) {
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < UPT; ++i) s = field_sum<NBITS>(v[i], s);
        s = wave_sum(s);
        if (lane == 0) zpart[par][wv] = s;
        __syncthreads();
        const double z = static_cast<double>(zpart[par][0] + zpart[par][1] + zpart[par][2] + zpart[par][3]);
        par ^= 1;
        uint32_t cur[UPT];
#pragma unroll
        for (int i = 0; i < UPT; ++i) cur[i] = v[i];
        load_row(t + gridDim.x, v);  // next row in flight while this one is written
        auto write_row = [&](auto clamp) {
#pragma unroll
            for (int i = 0; i < UPT; ++i) {
                const int64_t u = tid + 256 * i;
                if (u >= nunit) break;
                uint32_t o[PER], res[PER];
                spp::expand_samples<NBITS, BIG>(cur[i], o);
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    res[j] = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double x = static_cast<double>((o[j] >> (8 * k)) & 0xffu);
                        double y = fma(-z, static_cast<double>(wf[i][4 * j + k]), x) +
                                   static_cast<double>(bf[i][4 * j + k]);
                        if constexpr (decltype(clamp)::value) y = fmin(fmax(y, -2147483648.0), 2147483647.0);
                        res[j] |= (static_cast<uint32_t>(static_cast<int32_t>(y)) & 0xffu) << (8 * k);
                    }
                }
                uint8_t* dst = out + t * nchans + u * CPW;  // CPW bytes, CPW-aligned when vec
                if (vec && PER >= 4) {
#pragma unroll
                    for (int j = 0; j < PER; j += 4)
                        reinterpret_cast<uint4*>(dst)[j / 4] = make_uint4(res[j], res[(j + 1) % PER], res[(j + 2) % PER], res[(j + 3) % PER]);
                } else if (vec && PER == 2) {
                    *reinterpret_cast<uint2*>(dst) = make_uint2(res[0], res[1 % PER]);
                } else {
#pragma unroll
                    for (int j = 0; j < PER; ++j) reinterpret_cast<uint32_t*>(dst)[j] = res[j];
                }
            }
        };
        if (direct) write_row(std::false_type{});
        else write_row(std::true_type{});
    }
}

int check_nbits(int nbits) {
    SPP_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 32,
                "nbits must be 1, 2, 4, 8 or 32, got %d", nbits);
    return SPP_OK;
}

}  // namespace

extern "C" int spp_bandpass_workspace_size(int nbits, int64_t nchans, int64_t nsamps, int64_t* bytes) {
    spp::clear_error();
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(bytes != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0, "nchans must be > 0 and nsamps >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    *bytes = 0;
    if (nbits == 32 || nsamps == 0) return SPP_OK;
    const BpGeom g = bp_geom(nbits, nchans * nbits / 8, nsamps, true, true, true);
    *bytes = g.nchunks * nchans * static_cast<int64_t>(sizeof(uint32_t));
    return SPP_OK;
}

extern "C" int spp_bandpass(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                            float* out, int parallel, void* workspace, int64_t workspace_bytes,
                            void* stream) {
    spp::clear_error();
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(in != nullptr && out != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0, "nchans must be > 0 and nsamps >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(nsamps < (int64_t(1) << 31), "nsamps must be < 2^31");
    if (nsamps == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
    if (nbits == 32) {
        const dim3 grid(static_cast<unsigned>(spp::ceil_div(nchans, int64_t(256))));
        if (parallel)
            hipLaunchKernelGGL((k_bandpass_f32<false>), grid, dim3(256), 0, st, static_cast<const float*>(in),
                               nchans, nsamps, out);
        else
            hipLaunchKernelGGL((k_bandpass_f32<true>), grid, dim3(256), 0, st, static_cast<const float*>(in),
                               nchans, nsamps, out);
        SPP_LAUNCHED("k_bandpass_f32");
        return SPP_OK;
    }
    const uintptr_t a = reinterpret_cast<uintptr_t>(in);
    const BpGeom g = bp_geom(nbits, row_bytes, nsamps, a % 16 == 0, a % 8 == 0, a % 4 == 0);
    const int64_t need = g.nchunks * nchans * static_cast<int64_t>(sizeof(uint32_t));
    SPP_REQUIRE(workspace != nullptr, "null workspace");
    SPP_REQUIRE(workspace_bytes >= need, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(need));
    SPP_REQUIRE(g.gx <= 0x7fffffffLL && g.nchunks <= 65535, "gulp too large");
    auto* part = static_cast<uint32_t*>(workspace);
    const auto* src = static_cast<const uint8_t*>(in);
    const dim3 g1(static_cast<unsigned>(g.gx), static_cast<unsigned>(g.nchunks));
    const dim3 g2(static_cast<unsigned>(spp::ceil_div(nchans, int64_t(64))));
#define SPP_BP_PART(NB, BG, V)                                                                         \
    hipLaunchKernelGGL((k_bp_part<NB, BG, V>), g1, dim3(256), 0, st, src, nsamps, row_bytes, g.chunk_rows, \
                       nchans, part)
#define SPP_BP(NB, BG)                                                                                   \
    do {                                                                                                 \
        switch (g.vw) {                                                                                  \
            case 4: SPP_BP_PART(NB, BG, 4); break;                                                       \
            case 2: SPP_BP_PART(NB, BG, 2); break;                                                       \
            case 1: SPP_BP_PART(NB, BG, 1); break;                                                       \
            default: SPP_BP_PART(NB, BG, 0); break;                                                      \
        }                                                                                                \
        SPP_LAUNCHED("k_bp_part");                                                                       \
        if (parallel)                                                                                    \
            hipLaunchKernelGGL((k_bp_finish<NB, BG, false>), g2, dim3(64 * BPF_WAVES), 0, st, part,      \
                               g.nchunks, g.chunk_rows, nchans, src, nsamps, row_bytes, out);            \
        else                                                                                             \
            hipLaunchKernelGGL((k_bp_finish<NB, BG, true>), g2, dim3(64 * BPF_WAVES), 0, st, part,       \
                               g.nchunks, g.chunk_rows, nchans, src, nsamps, row_bytes, out);            \
    } while (0)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_BP(1, false); break;
        case 3: SPP_BP(1, true); break;
        case 4: SPP_BP(2, false); break;
        case 5: SPP_BP(2, true); break;
        case 8: SPP_BP(4, false); break;
        case 9: SPP_BP(4, true); break;
        default: SPP_BP(8, false); break;
    }
#undef SPP_BP
#undef SPP_BP_PART
    SPP_LAUNCHED("k_bp_finish");
    return SPP_OK;
}

extern "C" int spp_collapse(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                            float* out, int64_t out_start, void* stream) {
    spp::clear_error();
    (void)big;  // a sum over every field of a row does not depend on the bit order
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(in != nullptr && out != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0 && out_start >= 0, "nchans must be > 0, nsamps and out_start >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(nbits == 32 || nchans < (int64_t(1) << 24), "nchans must be < 2^24");
    if (nsamps == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
    float* o = out + out_start;
    if (nbits == 32) {
        hipLaunchKernelGGL(k_collapse_f32, dim3(static_cast<unsigned>(spp::ceil_div(nsamps, int64_t(256)))),
                           dim3(256), 0, st, static_cast<const float*>(in), nsamps, nchans, o);
        SPP_LAUNCHED("k_collapse_f32");
        return SPP_OK;
    }
    const uintptr_t a = reinterpret_cast<uintptr_t>(in);
    const int ub = (row_bytes % 16 == 0 && a % 16 == 0) ? 16 : (row_bytes % 4 == 0 && a % 4 == 0) ? 4 : 1;
    const int64_t blocks = spp::ceil_div(nsamps, int64_t(4 * CL_ROWS));
    SPP_REQUIRE(blocks <= 0x7fffffffLL, "gulp too long");
    const auto* src = static_cast<const uint8_t*>(in);
#define SPP_CL(NB)                                                                                      \
    do {                                                                                                \
        if (ub == 16) hipLaunchKernelGGL((k_collapse<NB, 16>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, src, nsamps, row_bytes, o); \
        else if (ub == 4) hipLaunchKernelGGL((k_collapse<NB, 4>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, src, nsamps, row_bytes, o); \
        else hipLaunchKernelGGL((k_collapse<NB, 1>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, src, nsamps, row_bytes, o); \
    } while (0)
    switch (nbits) {
        case 1: SPP_CL(1); break;
        case 2: SPP_CL(2); break;
        case 4: SPP_CL(4); break;
        default: SPP_CL(8); break;
    }
#undef SPP_CL
    SPP_LAUNCHED("k_collapse");
    return SPP_OK;
}

extern "C" int spp_mask_bytes(uint8_t* data, int64_t nbytes, int64_t row_bytes, const uint8_t* keep,
                              const uint8_t* set, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(data != nullptr && keep != nullptr && set != nullptr, "null argument");
    SPP_REQUIRE(row_bytes > 0 && nbytes >= 0 && nbytes % row_bytes == 0,
                "nbytes (%lld) must be a multiple of row_bytes (%lld)", static_cast<long long>(nbytes),
                static_cast<long long>(row_bytes));
    if (nbytes == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const uintptr_t align = reinterpret_cast<uintptr_t>(data) | reinterpret_cast<uintptr_t>(keep) |
                            reinterpret_cast<uintptr_t>(set);
    const int ub = (row_bytes % 16 == 0 && align % 16 == 0) ? 16 : (row_bytes % 4 == 0 && align % 4 == 0) ? 4 : 0;
    if (ub) {
        const int64_t nrows = nbytes / row_bytes;
        const int64_t gx = spp::ceil_div(row_bytes / ub, int64_t(256));
        const int64_t gy = std::max<int64_t>(1, std::min<int64_t>({nrows, spp::ceil_div(2048, gx), 65535}));
        SPP_REQUIRE(gx <= 0x7fffffffLL, "rows too long");
        const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
        if (ub == 16) hipLaunchKernelGGL((k_mask_rows<16>), grid, dim3(256), 0, st, data, nrows, row_bytes, keep, set);
        else hipLaunchKernelGGL((k_mask_rows<4>), grid, dim3(256), 0, st, data, nrows, row_bytes, keep, set);
        SPP_LAUNCHED("k_mask_rows");
        return SPP_OK;
    }
    const int64_t blocks = std::min<int64_t>(spp::ceil_div(nbytes, int64_t(256)), 65536);
    hipLaunchKernelGGL(k_mask_bytes, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, data, nbytes,
                       row_bytes, keep, set);
    SPP_LAUNCHED("k_mask_bytes");
    return SPP_OK;
}

extern "C" int spp_remove_zerodm(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                                 const float* chanwts, const float* bpass, void* out, void* stream) {
    spp::clear_error();
    if (check_nbits(nbits)) return SPP_EINVAL;
    SPP_REQUIRE(in != nullptr && chanwts != nullptr && bpass != nullptr && out != nullptr, "null argument");
    SPP_REQUIRE(nchans > 0 && nsamps >= 0, "nchans must be > 0 and nsamps >= 0");
    SPP_REQUIRE((nchans * nbits) % 8 == 0, "nchans * nbits must be a multiple of 8");
    SPP_REQUIRE(nsamps < (int64_t(1) << 31), "nsamps must be < 2^31");
    if (nsamps == 0) return SPP_OK;
    hipStream_t st = spp::as_stream(stream);
    const int64_t row_bytes = nchans * nbits / 8;
    const uintptr_t align = reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out);
    int64_t upt = spp::ceil_div(row_bytes / 4, int64_t(256));
    upt = upt <= 1 ? 1 : upt <= 2 ? 2 : upt <= 4 ? 4 : upt <= 8 ? 8 : upt;
    const int64_t cpw = nbits == 32 ? 1 : 32 / nbits;  // channels per dword
    if (nbits != 32 && row_bytes % 4 == 0 && align % 4 == 0 && upt * cpw <= 32) {
        const dim3 grid(static_cast<unsigned>(std::min<int64_t>(nsamps, 4 * spp::device_cu_count())));
        const auto* src = static_cast<const uint8_t*>(in);
        auto* dst = static_cast<uint8_t*>(out);
        // vector stores of a thread's 4 * 8/nbits output bytes when they are that aligned (rows
        // of nchans bytes keep it: 4 * 8/nbits divides nchans here)
        const int vec = reinterpret_cast<uintptr_t>(out) % static_cast<uintptr_t>(std::min<int64_t>(16, cpw)) == 0 ? 1 : 0;
#define SPP_ZDW(NB, BG, U) \
    hipLaunchKernelGGL((k_zerodm_w<NB, BG, U>), grid, dim3(256), 0, st, src, nsamps, nchans, row_bytes, chanwts, bpass, dst, vec)
#define SPP_ZDW_U(NB, BG)                                              \
    do {                                                               \
        if constexpr (32 / NB >= 32) SPP_ZDW(NB, BG, 1);               \
        else if constexpr (32 / NB >= 16) {                            \
            if (upt == 1) SPP_ZDW(NB, BG, 1); else SPP_ZDW(NB, BG, 2); \
        } else if constexpr (32 / NB >= 8) {                           \
            if (upt == 1) SPP_ZDW(NB, BG, 1);                          \
            else if (upt == 2) SPP_ZDW(NB, BG, 2);                     \
            else SPP_ZDW(NB, BG, 4);                                   \
        } else {                                                       \
            if (upt == 1) SPP_ZDW(NB, BG, 1);                          \
            else if (upt == 2) SPP_ZDW(NB, BG, 2);                     \
            else if (upt == 4) SPP_ZDW(NB, BG, 4);                     \
            else SPP_ZDW(NB, BG, 8);                                   \
        }                                                              \
    } while (0)
        switch (nbits * 2 + (big ? 1 : 0)) {
            case 2: SPP_ZDW_U(1, false); break;
            case 3: SPP_ZDW_U(1, true); break;
            case 4: SPP_ZDW_U(2, false); break;
            case 5: SPP_ZDW_U(2, true); break;
            case 8: SPP_ZDW_U(4, false); break;
            case 9: SPP_ZDW_U(4, true); break;
            default: SPP_ZDW_U(8, false); break;
        }
#undef SPP_ZDW_U
#undef SPP_ZDW
        SPP_LAUNCHED("k_zerodm_w");
        return SPP_OK;
    }
    const dim3 grid(static_cast<unsigned>(nsamps));
#define SPP_ZD(NB, BG) \
    hipLaunchKernelGGL((k_zerodm<NB, BG>), grid, dim3(256), 0, st, in, nchans, row_bytes, chanwts, bpass, out)
    switch (nbits * 2 + (big ? 1 : 0)) {
        case 2: SPP_ZD(1, false); break;
        case 3: SPP_ZD(1, true); break;
        case 4: SPP_ZD(2, false); break;
        case 5: SPP_ZD(2, true); break;
        case 8: SPP_ZD(4, false); break;
        case 9: SPP_ZD(4, true); break;
        case 16: case 17: SPP_ZD(8, false); break;
        default: SPP_ZD(32, false); break;
    }
#undef SPP_ZD
    SPP_LAUNCHED("k_zerodm");
    return SPP_OK;
}
