// Matched-filter stage after the sweep: running normalisation, boxcar matched filter
// with a per-row argmax, and the general template-bank convolution.
//
// Everything here is a small fraction of the sweep's cost (a few HBM passes over the
// plane). Reductions use fixed orders, so results are deterministic run to run.
//
// Running normalisation (one DM series x of length n, window W):
//   p = x padded symmetrically by (W/2, W/2 - (W even)) samples (the padding of
//   core/stats.py:231-236), d = p - x[0]. For output t the window sums are
//   S1(t) = sum_{k=t}^{t+W-1} d[k] and S2(t) likewise for d^2 (float64),
//   m = S1/W, v = max(S2/W - m*m, 0), s = sqrt(v) (s == 0 -> 1),
//   z[t] = float32((x[t] - (m + x[0])) / s)  (1/s by a refined reciprocal square root).
// One block covers RN_CHB outputs of a row: it sums the first window, then slides it,
// S(t + 1) = S(t) + d[t + W] - d[t], one block-wide scan of those differences per 2048
// outputs. For integer-valued series (the sweep's planes) every partial sum is an exact
// float64 integer.
//
// Boxcar MF (core/kernels.py:861-906 closed form, row length n, Ng = good size >= n):
//   zt[i] = z[(i mod Ng) mod n], S = sum_{i<Ng} zt[i],
//   conv_w[t] = (sum_{j<w} zt[t+j] - w*S/Ng) / sqrt(w*(1 - w/Ng)),  t < n,
//   argmax over [width][t] in template-major order with the first maximum kept.
// That is the reference's result when Ng is even. For an odd Ng the reference's irfft has
// Ng - 1 output samples and the row goes through the FFT path of spp_fft.hpp instead; a
// row of exactly Ng (odd) samples is refused, as the reference cannot fill it either.
#include "spp_fft.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace {

constexpr int RN_SEG = 2048;     // padded samples per segment / outputs per k_norm block
constexpr int RN_THREADS = 256;
constexpr int RN_PER = RN_SEG / RN_THREADS;  // 8
constexpr int RN_CHB = 16 * RN_SEG;  // outputs per normalisation block
constexpr int MF_B = 2048;       // outputs per MF block
constexpr int MF_THREADS = 256;
constexpr int MF_WMAX = 1024;    // largest boxcar width supported
constexpr int MF_MAXW = 32;      // largest number of widths
constexpr int MF_NSP = 32;       // partial sums per row for the circular total
constexpr int CT_B = 1024;       // outputs per template-convolution block
constexpr int CT_TAPS = 2048;    // largest template length

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

// i mod m for 0 <= i < k*m with small k, by subtraction (64-bit division is a long
// instruction sequence on the GPU).
__device__ __forceinline__ int64_t wrap_index(int64_t i, int64_t m) {
    while (i >= m) i -= m;
    return i;
}

__device__ __forceinline__ int64_t sym_index(int64_t i, int64_t n) {
    // numpy 'symmetric' extension: period 2n, mirrored including the edge sample.
    if (i >= 0 && i < n) return i;  // interior: no reflection (and no 64-bit division)
    const int64_t two_n = 2 * n;
    int64_t m = i % two_n;
    if (m < 0) m += two_n;
    return (m < n) ? m : (two_n - 1 - m);
}

__device__ __forceinline__ double wave_incl_scan(double v);
__device__ __forceinline__ double lane63(double v);

// Block-wide sum of a double in a fixed order (the wave's inclusive scan, then the wave
// totals in wave order). All threads get the result. blockDim.x is a multiple of 64 and all
// threads take part.
__device__ double block_sum(double v, double* red) {
    v = lane63(wave_incl_scan(v));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) r += red[w];
    __syncthreads();
    return r;
}

// A double moved between lanes by one DPP control (both 32-bit halves); lanes whose source
// is outside the row, or whose row is masked off, get 0.0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Inclusive scan of a double across one wave (fixed order): within each row of 16 lanes by
// row shifts of 1, 2, 4, 8, then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then lane 31's running total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ double wave_incl_scan(double v) {
    v += dpp_move<0x111, 0xf>(v);  // row_shr:1
    v += dpp_move<0x112, 0xf>(v);  // row_shr:2
    v += dpp_move<0x114, 0xf>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf>(v);  // row_shr:8
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    return v;
}

// Lane 63's value, in every lane.
__device__ __forceinline__ double lane63(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), 63);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// The smallest int across the wave, in every lane (row shifts and broadcasts as in
// wave_incl_scan; lanes without a source keep INT32_MAX).
__device__ __forceinline__ int32_t wave_min(int32_t v) {
    constexpr int kMax = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x111, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x112, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x114, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x118, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x142, 0xa, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// The largest unsigned int across the wave, in every lane.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
    v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// The value of the previous lane (0.0 in lane 0): wave_shr:1.
__device__ __forceinline__ double wave_prev(double v) { return dpp_move<0x138, 0xf>(v); }

// As dpp_move, but lanes without a source keep -inf.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move_or_neginf(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(static_cast<int>(0xfff00000u), static_cast<int>(b >> 32), CTRL,
                                               ROW_MASK, 0xf, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// The largest value across the wave (NaN lanes ignored), the same in every lane: the scan
// steps of wave_incl_scan with max, then lane 63's result. All 64 lanes must be active.
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_move_or_neginf<0x111, 0xf>(v));
    v = fmax(v, dpp_move_or_neginf<0x112, 0xf>(v));
    v = fmax(v, dpp_move_or_neginf<0x114, 0xf>(v));
    v = fmax(v, dpp_move_or_neginf<0x118, 0xf>(v));
    v = fmax(v, dpp_move_or_neginf<0x142, 0xa>(v));
    v = fmax(v, dpp_move_or_neginf<0x143, 0xc>(v));
    return lane63(v);
}

// Exclusive block scan of two doubles per thread, also returning the block totals (ta, tb):
// within each wave by shuffles, then the waves' totals from LDS in wave order. Two barriers;
// sa/sb hold one entry per wave.
__device__ void block_exscan2_total(double& a, double& b, double& ta, double& tb, double* sa, double* sb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double ia = wave_incl_scan(a), ib = wave_incl_scan(b);
    const double xa = wave_prev(ia), xb = wave_prev(ib);
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    double pa = 0.0, pb = 0.0, qa = 0.0, qb = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        if (w < wave) {
            pa += sa[w];
            pb += sb[w];
        }
        qa += sa[w];
        qb += sb[w];
    }
    __syncthreads();
    a = pa + xa;
    b = pb + xb;
    ta = qa;
    tb = qb;
}

// Exclusive block scan of two doubles per thread: within each wave by shuffles, then the
// preceding waves' totals from LDS. Two barriers; sa/sb hold one entry per wave.
__device__ void block_exscan2(double& a, double& b, double* sa, double* sb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double ia = wave_incl_scan(a), ib = wave_incl_scan(b);
    const double xa = wave_prev(ia), xb = wave_prev(ib);
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    double pa = 0.0, pb = 0.0;
    for (int w = 0; w < wave; ++w) {
        pa += sa[w];
        pb += sb[w];
    }
    __syncthreads();
    a = pa + xa;
    b = pb + xb;
}

// Exclusive block scan of one double per thread, also returning the block total. Two
// barriers; sa holds one entry per wave.
__device__ void block_exscan1_total(double& a, double& ta, double* sa) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double ia = wave_incl_scan(a);
    const double xa = wave_prev(ia);
    if (lane == 63) sa[wave] = ia;
    __syncthreads();
    double pa = 0.0, qa = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
        if (w < wave) pa += sa[w];
        qa += sa[w];
    }
    __syncthreads();
    a = pa + xa;
    ta = qa;
}

// ---------------------------------------------------------------- running normalise

// z of one sample x from its window sums s1 = sum d and s2 = sum d^2 (d = p - ref, W
// samples): m = s1/W, v = max(s2/W - m^2, 0), z = (x - (m + ref)) / sqrt(v), with 1 for
// sqrt(v) where v == 0. 1/sqrt(v) is v_rsq_f64 (about 2^-23 relative) refined by one Newton
// step, r <- r + r (1 - v r^2) / 2, to about 2^-45, so the float32 z equals the correctly
// rounded quotient except where that quotient lies within ~2^-45 of a float32 rounding
// boundary (then it is the neighbouring float32, one ulp away).
__device__ __forceinline__ float znorm(float x, double s1, double s2, double invw, double ref) {
    const double m = s1 * invw;
    const double v = s2 * invw - m * m;
    // branch-free: the refinement runs on every lane (inf or NaN where v <= 0, then replaced)
    const double hv = 0.5 * v;
    double r = __builtin_amdgcn_rsq(v);
    r = fma(r, fma(-hv * r, r, 0.5), r);
    r = (v > 0.0) ? r : 1.0;
    return static_cast<float>((static_cast<double>(x) - (m + ref)) * r);
}

struct RowDesc {
    int64_t n;  // row length
};

// z for outputs [blk * RN_CHB, ...) of one row (blocks numbered so that one XCD takes
// consecutive blocks of a row, which share the window's samples in its L2).
__global__ __launch_bounds__(RN_THREADS) void k_norm(const float* __restrict__ x, int64_t ldx,
                                                     const RowDesc* __restrict__ rows, int64_t nb_row,
                                                     int64_t nblocks, int32_t window,
                                                     float* __restrict__ z, int64_t ldz) {
    __shared__ double sa[RN_THREADS / 64], sb[RN_THREADS / 64], red[RN_THREADS / 64];
    const int64_t b = blockIdx.x;
    const int64_t logical = (b % 8) * (gridDim.x / 8) + b / 8;
    if (logical >= nblocks) return;
    const int64_t row = logical / nb_row;
    const int64_t n = rows[row].n;
    const int64_t t0 = (logical % nb_row) * RN_CHB;
    if (t0 >= n) return;
    const int64_t tend = min<int64_t>(t0 + RN_CHB, n);
    const int tid = threadIdx.x;
    const float* __restrict__ xr = x + row * ldx;
    float* __restrict__ zr = z + row * ldz;
    const int64_t L = window / 2;
    const double ref = xr[0];
    auto dpad = [&](int64_t k) { return static_cast<double>(xr[sym_index(k - L, n)]) - ref; };
    // A chunk's samples for the thread's outputs t = u0 + RN_PER*tid + j: p[t], p[t + W] and
    // x[t]; the next chunk's are loaded while the current one is scanned.
    float c0[RN_PER], cn[RN_PER], cx[RN_PER];
    auto load_chunk = [&](int64_t u0, float (&v0)[RN_PER], float (&vn)[RN_PER], float (&vx)[RN_PER]) {
        if (u0 >= L && u0 + RN_SEG + window - L <= n && u0 + RN_SEG <= tend) {
            // interior chunk (almost all of them): no reflection, no bounds tests
            const float* __restrict__ q = xr + u0 + RN_PER * tid;
#pragma unroll
            for (int j = 0; j < RN_PER; ++j) {
                v0[j] = q[j - L];
                vn[j] = q[j + window - L];
                vx[j] = q[j];
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            const int64_t t = u0 + RN_PER * tid + j;
            v0[j] = vn[j] = vx[j] = 0.f;
            if (t < tend) {
                v0[j] = xr[sym_index(t - L, n)];
                vn[j] = xr[sym_index(t + window - L, n)];
                vx[j] = xr[t];
            }
        }
    };
    load_chunk(t0, c0, cn, cx);
    // Window sums at t0: padded indices [t0, t0 + W).
    double s1 = 0.0, s2 = 0.0;
    if (t0 >= L && t0 + window - L <= n) {
        for (int64_t i = tid; i < window; i += RN_THREADS) {
            const double d = static_cast<double>(xr[t0 + i - L]) - ref;
            s1 += d;
            s2 += d * d;
        }
    } else {
        for (int64_t i = tid; i < window; i += RN_THREADS) {
            const double d = dpad(t0 + i);
            s1 += d;
            s2 += d * d;
        }
    }
    double S1 = block_sum(s1, red), S2 = block_sum(s2, red);
    const double invw = 1.0 / static_cast<double>(window);
    for (int64_t u0 = t0; u0 < tend; u0 += RN_SEG) {
        float f0[RN_PER] = {}, fn[RN_PER] = {}, fx[RN_PER] = {};
        if (u0 + RN_SEG < tend) load_chunk(u0 + RN_SEG, f0, fn, fx);
        // Thread's outputs u0 + RN_PER*tid + j: local exclusive prefix of d[t+W] - d[t]
        // (zero past tend).
        double p1[RN_PER], p2[RN_PER];
        double a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            // samples past tend were loaded as 0: both differences are exactly 0 there
            p1[j] = a1;
            p2[j] = a2;
            const double dn = static_cast<double>(cn[j]) - ref, d0 = static_cast<double>(c0[j]) - ref;
            a1 += dn - d0;
            a2 += dn * dn - d0 * d0;
        }
        double e1 = a1, e2 = a2, T1, T2;
        block_exscan2_total(e1, e2, T1, T2, sa, sb);
        float zv[RN_PER];
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) zv[j] = znorm(cx[j], S1 + e1 + p1[j], S2 + e2 + p2[j], invw, ref);
        float* __restrict__ zt = zr + u0 + RN_PER * tid;
        if (u0 + RN_SEG <= tend) {
#pragma unroll
            for (int j = 0; j < RN_PER; ++j) zt[j] = zv[j];
        } else {
            const int64_t tl = tend - (u0 + RN_PER * tid);
#pragma unroll
            for (int j = 0; j < RN_PER; ++j)
                if (j < tl) zt[j] = zv[j];
        }
        S1 += T1;
        S2 += T2;
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            c0[j] = f0[j];
            cn[j] = fn[j];
            cx[j] = fx[j];
        }
    }
}

// ---------------------------------------------------------------- boxcar MF

struct MfRow {
    int64_t n;
    int64_t ng;
};

struct Best {
    float val;
    int32_t w;
    int64_t t;
};

__device__ __forceinline__ bool better(float v, int32_t w, int64_t t, const Best& b) {
    if (v > b.val) return true;
    if (v < b.val || v != v) return false;
    return (w < b.w) || (w == b.w && t < b.t);
}

// Partial sums of zt over [c * MF_SP, (c + 1) * MF_SP) of [0, Ng) per row (part c of
// MF_NSP); k_boxcar_part adds a row's parts in order, so S = sum_{i<Ng} zt[i] is
// deterministic.
__global__ __launch_bounds__(MF_THREADS) void k_circ_sum(const float* __restrict__ z, int64_t ldz,
                                                         const MfRow* __restrict__ rows,
                                                         double* __restrict__ Sp) {
    __shared__ double red[MF_THREADS / 64];
    const int64_t row = blockIdx.y;
    const MfRow r = rows[row];
    if (r.ng & 1) return;  // odd good size: the FFT path (spp_fft.hpp) handles the row
    const float* __restrict__ zr = z + row * ldz;
    const int64_t per = (r.ng + MF_NSP - 1) / MF_NSP;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * per, i1 = min(i0 + per, r.ng);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += MF_THREADS) s += static_cast<double>(zr[wrap_index(i, r.n)]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) Sp[row * MF_NSP + blockIdx.x] = s;
}

__global__ __launch_bounds__(MF_THREADS) void k_boxcar_part(
    const float* __restrict__ z, int64_t ldz, const MfRow* __restrict__ rows,
    const double* __restrict__ S, const int32_t* __restrict__ widths, int nw, int wmax,
    int64_t chunk_pitch, Best* __restrict__ part) {
    __shared__ double Q[MF_B + MF_WMAX + 1];
    __shared__ double sa[MF_THREADS], sb[MF_THREADS];
    __shared__ double coef_mean[MF_MAXW], coef_inv[MF_MAXW];
    __shared__ int32_t sw[MF_MAXW];
    __shared__ Best bred[MF_THREADS];
    const int64_t row = blockIdx.y;
    const MfRow r = rows[row];
    const int64_t a = static_cast<int64_t>(blockIdx.x) * MF_B;
    const int tid = threadIdx.x;
    Best best;
    best.val = -FLT_MAX;
    best.w = 0x7fffffff;
    best.t = INT64_MAX;
    if (a >= r.n || (r.ng & 1)) {  // odd good size: partials come from k_conv_argmax_part
        if (tid == 0) part[row * chunk_pitch + blockIdx.x] = best;
        return;
    }
    const float* __restrict__ zr = z + row * ldz;
    double srow = 0.0;
    for (int i = 0; i < MF_NSP; ++i) srow += S[row * MF_NSP + i];
    const double ng = static_cast<double>(r.ng);
    if (tid < nw) {
        const double w = widths[tid];
        coef_mean[tid] = w * srow / ng;
        coef_inv[tid] = 1.0 / sqrt(w * (1.0 - w / ng));
        sw[tid] = widths[tid];
    }
    const int64_t lw = min<int64_t>(MF_B, r.n - a);
    const int nel = static_cast<int>(lw) + wmax;  // zt[a .. a + nel)
    // Each thread scans KE consecutive values (KE odd: its LDS reads and writes hit distinct
    // banks). The values are first staged in Q's storage as floats with coalesced loads.
    constexpr int KE = 13;
    static_assert(KE * MF_THREADS >= MF_B + MF_WMAX && KE % 2 == 1, "odd run covering a block");
    static_assert(KE * MF_THREADS * sizeof(float) <= sizeof(Q), "staging fits in Q");
    float* Zs = reinterpret_cast<float*>(Q);
    float zv[KE];
    if (a + nel <= r.n) {
        // interior block: all KE loads in flight at once
#pragma unroll
        for (int e = 0; e < KE; ++e) {
            const int k = tid + MF_THREADS * e;
            zv[e] = (k < nel) ? zr[a + k] : 0.f;
        }
    } else {
#pragma unroll
        for (int e = 0; e < KE; ++e) {
            const int k = tid + MF_THREADS * e;
            zv[e] = 0.f;
            if (k < nel) {
                const int64_t i = wrap_index(a + k, r.ng);  // a + k < n + MF_B + MF_WMAX
                zv[e] = zr[wrap_index(i, r.n)];             // ng < 2n
            }
        }
    }
#pragma unroll
    for (int e = 0; e < KE; ++e) Zs[tid + MF_THREADS * e] = zv[e];
    __syncthreads();
    double loc[KE];
    double tot = 0.0;
#pragma unroll
    for (int e = 0; e < KE; ++e) {
        tot += static_cast<double>(Zs[tid * KE + e]);
        loc[e] = tot;
    }
    __syncthreads();  // Q is written below over the staged floats
    double ex = tot, dummy = 0.0;
    block_exscan2(ex, dummy, sa, sb);
    if (tid == 0) Q[0] = 0.0;
#pragma unroll
    for (int e = 0; e < KE; ++e) {
        const int k = tid * KE + e;
        if (k < nel) Q[k + 1] = ex + loc[e];
    }
    __syncthreads();
    // Widths outer (their constants are uniform), the thread's outputs u = tid + 256i inner,
    // with Q[u] held in registers. t increases within a thread, so a strict > keeps each
    // width's first maximum; better() then applies the template-major tie order.
    constexpr int NU = MF_B / MF_THREADS;
    double q0[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int u = tid + MF_THREADS * i;
        q0[i] = (u < lw) ? Q[u] : 0.0;
    }
    for (int wi = 0; wi < nw; ++wi) {
        const int w = sw[wi];
        const double cm = coef_mean[wi], ci = coef_inv[wi];
        float bv = -FLT_MAX;
        int bt = -1;
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            const int u = tid + MF_THREADS * i;
            if (u < lw) {
                const float cv = static_cast<float>((Q[u + w] - q0[i] - cm) * ci);
                if (cv > bv) {
                    bv = cv;
                    bt = u;
                }
            }
        }
        if (bt >= 0 && better(bv, wi, a + bt, best)) {
            best.val = bv;
            best.w = wi;
            best.t = a + bt;
        }
    }
    bred[tid] = best;
    __syncthreads();
    for (int s = MF_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const Best o = bred[tid + s];
            if (better(o.val, o.w, o.t, bred[tid])) bred[tid] = o;
        }
        __syncthreads();
    }
    if (tid == 0) part[row * chunk_pitch + blockIdx.x] = bred[0];
}

__global__ void k_boxcar_final(const Best* __restrict__ part, int64_t nrows, int64_t chunk_pitch,
                               const MfRow* __restrict__ rows, float* __restrict__ snr,
                               int32_t* __restrict__ peak, int32_t* __restrict__ widx) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const int64_t nch = (rows[row].n + MF_B - 1) / MF_B;
    Best b = part[row * chunk_pitch];
    for (int64_t c = 1; c < nch; ++c) {
        const Best o = part[row * chunk_pitch + c];
        if (better(o.val, o.w, o.t, b)) b = o;
    }
    snr[row] = b.val;
    peak[row] = static_cast<int32_t>(b.t);
    widx[row] = b.w;
}

// ---------------------------------------------------------------- fused normalise + boxcar MF
//
// The running normalisation and the boxcar MF of the rows of a plane x in one pass: z never
// leaves the block. A block takes the RN_CHB outputs of one row that k_norm's block takes, and
// normalises them RN_SEG at a time with the same window sums and arithmetic (so z is the
// float32 k_norm writes). Each segment's z is prefix-summed in float64, relative to the block
// start, into an LDS ring P of two segments; the box sums of the previous segment are then
// P[u + w] - P[u]. A box reaches wmax - 1 <= MF_WMAX - 1 < RN_SEG samples into the next
// segment, so the block normalises one segment past its chunk. The circular series wraps at
// the row end: zt[t] = z[t - n] for n <= t < Ng and z[t - Ng] past Ng, both below wmax, so a
// block whose boxes cross the row end first normalises z[0, RN_SEG) into LDS.
//
// Per width the block keeps the first maximum of the box sum over its outputs, and the sum of
// its z (counted twice for t < Ng - n, the samples that recur in zt[n, Ng)) towards S. The
// maximum is found in two passes per segment: each lane's largest sum per width (straight
// reads and float64 max), then exact first-position tracking only in the waves where some
// lane reaches the largest sum the block has tracked so far -- about one segment in nine
// after the first, as a new running maximum is rare.
// k_nb_final adds a row's chunk sums in order, takes each width's first maximum over the
// chunks, forms conv_w = (box - w S / Ng) / sqrt(w (1 - w / Ng)) in float64, rounds to
// float32 and picks the template-major first maximum. The float32 rounding is monotone, so
// each width's maximum is k_boxcar_part's; the first position can differ from k_boxcar_part's
// only where two box sums that differ in float64 round to the same float32 value.

struct NbRow {
    int64_t n;
    int64_t ng;
    int64_t odd;  // nonzero: the row goes through the FFT path
};

constexpr int NB_RING = 2 * RN_SEG;
constexpr int NB_NONE = 0x3fffffff;   // position code of "no candidate" (codes are < 2^10)
constexpr int NB_NOPOS = 0x7fffffff;  // position of "no candidate" (positions are < 2^31 - 1)

// An unsigned key with the order of the doubles (NaN aside): sign bit set for x >= 0,
// all bits flipped for x < 0.
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// NWMAX: widths the per-width loops are unrolled for; EXACT: nw == NWMAX (no per-width
// branch, so the box reads of one width can be in flight while the previous one is reduced).
// Three blocks per CU (the LDS allows three); two for the 32-width form, whose per-width
// registers would not fit three.
template <int NWMAX, bool EXACT, bool ALIGNED>
__global__ __launch_bounds__(RN_THREADS, NWMAX > 12 ? 2 : 3) void k_norm_boxcar(
    const float* __restrict__ x, int64_t ldx, const NbRow* __restrict__ rows, int64_t nb_row,
    int64_t nblocks, int32_t window, const int32_t* __restrict__ widths, int nw, int32_t wmax,
    double* __restrict__ part_s, double* __restrict__ part_v, int32_t* __restrict__ part_t) {
    __shared__ double ring[NB_RING + MF_WMAX];
    __shared__ float head[RN_SEG];
    __shared__ double sa[RN_THREADS / 64], sb[RN_THREADS / 64], red[RN_THREADS / 64];
    // per width: the upper 32 bits of the order key (order_key) of the largest box sum any lane
    // of the block has tracked so far
    __shared__ uint32_t tau_hi[NWMAX];
    __shared__ double wv[NWMAX][RN_THREADS / 64];
    __shared__ int32_t wt[NWMAX][RN_THREADS / 64];
    const int64_t b = blockIdx.x;
    const int64_t logical = (b % 8) * (gridDim.x / 8) + b / 8;
    if (logical >= nblocks) return;
    const int64_t row = logical / nb_row;
    const int64_t chunk = logical % nb_row;
    const NbRow r = rows[row];
    const int64_t n = r.n;
    const int64_t t0 = chunk * RN_CHB;
    if (t0 >= n || r.odd) return;
    const int64_t tend = min<int64_t>(t0 + RN_CHB, n);  // the chunk's outputs
    const int64_t tn = min<int64_t>(tend + RN_SEG, n);  // samples normalised (one segment more)
    const int64_t wrapn = r.ng - n;                      // zt[n, Ng) = z[0, Ng - n)
    const int tid = threadIdx.x;
    const float* __restrict__ xr = x + row * ldx;
    const int64_t L = window / 2;
    const double ref = xr[0];
    const double invw = 1.0 / static_cast<double>(window);

    // Samples of a segment for the thread's outputs t = u0 + RN_PER*tid + j: p[t], p[t + W]
    // and x[t] (p = x reflected by L samples at both ends), zero from `lim` on.
    auto load_seg = [&](int64_t u0, int64_t lim, float (&v0)[RN_PER], float (&vn)[RN_PER], float (&vx)[RN_PER]) {
        if (u0 >= L && u0 + RN_SEG + window - L <= n && u0 + RN_SEG <= lim) {
            const float* __restrict__ q = xr + u0 + RN_PER * tid;
            if constexpr (ALIGNED) {
                typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
                for (int h = 0; h < RN_PER / 4; ++h) {
                    const f4 a0 = *reinterpret_cast<const f4*>(q + 4 * h - L);
                    const f4 an = *reinterpret_cast<const f4*>(q + 4 * h + window - L);
                    const f4 ax = *reinterpret_cast<const f4*>(q + 4 * h);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        v0[4 * h + e] = a0[e];
                        vn[4 * h + e] = an[e];
                        vx[4 * h + e] = ax[e];
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < RN_PER; ++j) {
                    v0[j] = q[j - L];
                    vn[j] = q[j + window - L];
                    vx[j] = q[j];
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            const int64_t t = u0 + RN_PER * tid + j;
            v0[j] = vn[j] = vx[j] = 0.f;
            if (t < lim) {
                v0[j] = xr[sym_index(t - L, n)];
                vn[j] = xr[sym_index(t + window - L, n)];
                vx[j] = xr[t];
            }
        }
    };
    // Window sums at u0 (padded indices [u0, u0 + W)), in k_norm's order.
    auto window_sums = [&](int64_t u0, double& S1, double& S2) {
        double s1 = 0.0, s2 = 0.0;
        if (u0 >= L && u0 + window - L <= n) {
            for (int64_t i = tid; i < window; i += RN_THREADS) {
                const double d = static_cast<double>(xr[u0 + i - L]) - ref;
                s1 += d;
                s2 += d * d;
            }
        } else {
            for (int64_t i = tid; i < window; i += RN_THREADS) {
                const double d = static_cast<double>(xr[sym_index(u0 + i - L, n)]) - ref;
                s1 += d;
                s2 += d * d;
            }
        }
        S1 = block_sum(s1, red);
        S2 = block_sum(s2, red);
    };
    // z of one segment (k_norm's arithmetic); advances the window sums by RN_SEG outputs.
    auto norm_seg = [&](int64_t u0, int64_t lim, const float (&v0)[RN_PER], const float (&vn)[RN_PER],
                        const float (&vx)[RN_PER], double& S1, double& S2, float (&z)[RN_PER]) {
        double p1[RN_PER], p2[RN_PER];
        double a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            // samples from lim on were loaded as 0: both differences are exactly 0 there
            p1[j] = a1;
            p2[j] = a2;
            const double dn = static_cast<double>(vn[j]) - ref, d0 = static_cast<double>(v0[j]) - ref;
            a1 += dn - d0;
            a2 += dn * dn - d0 * d0;
        }
        double e1 = a1, e2 = a2, T1, T2;
        block_exscan2_total(e1, e2, T1, T2, sa, sb);
        const int64_t tl = lim - (u0 + RN_PER * tid);  // the thread's outputs j < tl are inside
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            const float zj = znorm(vx[j], S1 + e1 + p1[j], S2 + e2 + p2[j], invw, ref);
            z[j] = (j < tl) ? zj : 0.f;
        }
        S1 += T1;
        S2 += T2;
    };

    // z[0, nhead) of the row when this block's boxes cross the row end.
    const int64_t nhead = min<int64_t>(n, RN_SEG);
    const bool need_head = tend + wmax > n;
    if (need_head) {
        float h0[RN_PER], hn[RN_PER], hx[RN_PER], hz[RN_PER];
        load_seg(0, nhead, h0, hn, hx);
        double H1, H2;
        window_sums(0, H1, H2);
        norm_seg(0, nhead, h0, hn, hx, H1, H2, hz);
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) head[RN_PER * tid + j] = hz[j];
    }

    float c0[RN_PER], cn[RN_PER], cx[RN_PER];
    load_seg(t0, tn, c0, cn, cx);
    double S1, S2;
    window_sums(t0, S1, S2);
    const int lane = tid & 63, wave = tid >> 6;
    if (tid < NWMAX) tau_hi[tid] = 0;  // below every key; first read after the scans' barriers
    double bestv[NWMAX];
    int32_t bestt[NWMAX];  // position codes (segment << 3) | i
#pragma unroll
    for (int wi = 0; wi < NWMAX; ++wi) {
        bestv[wi] = -INFINITY;
        bestt[wi] = NB_NONE;
    }
    // The widths (<= MF_WMAX < 2^11), three to a VGPR: held in vector registers for the whole
    // block (as scalars the compiler reloads them in the loop, and each reload's wait also
    // drains the LDS reads in flight).
    constexpr int NPACK = (NWMAX + 2) / 3;
    uint32_t wpack[NPACK];
#pragma unroll
    for (int p = 0; p < NPACK; ++p) {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (3 * p + j < NWMAX && 3 * p + j < nw) v |= static_cast<uint32_t>(widths[3 * p + j]) << (11 * j);
        asm volatile("v_mov_b32 %0, %1" : "=v"(wpack[p]) : "s"(v));
    }
    auto width_of = [&](int wi) { return static_cast<int>((wpack[wi / 3] >> (11 * (wi % 3))) & 0x7ffu); };
    double carry = 0.0;  // P at the current segment's start
    double s_acc = 0.0;  // the thread's share of the chunk's z sum
    const int nseg = static_cast<int>((tend - t0 + RN_SEG - 1) / RN_SEG);
    for (int k = 0; k <= nseg; ++k) {
        const int64_t u0 = t0 + static_cast<int64_t>(k) * RN_SEG;
        float z[RN_PER];
        if (u0 < tn) {
            norm_seg(u0, tn, c0, cn, cx, S1, S2, z);
        } else {
#pragma unroll
            for (int j = 0; j < RN_PER; ++j) z[j] = 0.f;
        }
        // the next segment's samples load while this one is summed and boxed
        if (k < nseg && u0 + RN_SEG < tn) load_seg(u0 + RN_SEG, tn, c0, cn, cx);
        if (need_head && u0 + RN_SEG > n) {  // circular samples past the row end (the first wmax are read)
#pragma unroll
            for (int j = 0; j < RN_PER; ++j) {
                const int64_t t = u0 + RN_PER * tid + j;
                if (t >= n) {
                    const int64_t i = (t < r.ng) ? t - n : t - r.ng;
                    z[j] = (i < nhead) ? head[i] : 0.f;
                }
            }
        }
        double pre[RN_PER];
        double tot = 0.0;
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            pre[j] = tot;
            tot += static_cast<double>(z[j]);
        }
        if (k < nseg) {
            if (u0 + RN_SEG <= tend && u0 >= wrapn) {
                s_acc += tot;
            } else {
#pragma unroll
                for (int j = 0; j < RN_PER; ++j) {
                    const int64_t t = u0 + RN_PER * tid + j;
                    if (t < tend) s_acc += (t < wrapn) ? 2.0 * static_cast<double>(z[j]) : static_cast<double>(z[j]);
                }
            }
        }
        double ex = tot, T;
        block_exscan1_total(ex, T, sa);
        {
            // slots [0, MF_WMAX) are mirrored past the ring's end, so that a box of the
            // second half reads its far end without wrapping the index
            const int base = (k * RN_SEG + RN_PER * tid) & (NB_RING - 1);
            const double p0 = carry + ex;
#pragma unroll
            for (int j = 0; j < RN_PER; ++j) ring[base + j] = p0 + pre[j];
            if (base < MF_WMAX) {
#pragma unroll
                for (int j = 0; j < RN_PER; ++j) ring[NB_RING + base + j] = p0 + pre[j];
            }
        }
        carry += T;
        __syncthreads();
        if (k >= 1) {
            // boxes of the previous segment: outputs a + tid + RN_THREADS*i, whose P[u] sit at
            // ring[b + RN_THREADS*i] and P[u + w] at ring[b + RN_THREADS*i + w]
            const int64_t a = t0 + static_cast<int64_t>(k - 1) * RN_SEG;
            const bool full = a + RN_SEG <= tend;
            const double* __restrict__ rb = ring + (((k - 1) * RN_SEG) & (NB_RING - 1)) + tid;
            double q0[RN_PER];
#pragma unroll
            for (int i = 0; i < RN_PER; ++i) q0[i] = rb[RN_THREADS * i];
            // a thread's outputs are coded (segment << 3) | i, relative to this segment here
            const int frame = (k - 1) << 3;
            // outputs past the chunk's end (last segment only) take no part
            int nvalid = RN_PER;
            if (!full) {
                const int64_t left = tend - a - tid;  // outputs i with RN_THREADS*i < left
                nvalid = left <= 0 ? 0 : static_cast<int>(min<int64_t>(RN_PER, (left + RN_THREADS - 1) / RN_THREADS));
            }
            auto boxes = [&](auto all_valid) {
                // pass 1: per width, whether some lane of the wave reaches the largest sum tracked
                // in the block so far (one bit per width)
                auto max_of = [&](const double (&v)[RN_PER]) {
                    double d[RN_PER];
#pragma unroll
                    for (int i = 0; i < RN_PER; ++i) {
                        d[i] = v[i] - q0[i];
                        if constexpr (!decltype(all_valid)::value) d[i] = (i < nvalid) ? d[i] : -INFINITY;
                    }
                    return fmax(fmax(fmax(d[0], d[1]), fmax(d[2], d[3])), fmax(fmax(d[4], d[5]), fmax(d[6], d[7])));
                };
                auto reaches = [&](int wi, double m) {
                    return __any(static_cast<uint32_t>(order_key(m) >> 32) >= tau_hi[wi]) ? 1u : 0u;
                };
                uint32_t gate = 0;
                if constexpr (EXACT) {
                    // software-pipelined by hand: the reads of width wi + 1 are in flight while
                    // width wi is reduced (the scheduling barriers keep the compiler from
                    // hoisting every width's reads at once and spilling them)
                    double buf[2][RN_PER];
#pragma unroll
                    for (int i = 0; i < RN_PER; ++i) buf[0][i] = rb[width_of(0) + RN_THREADS * i];
#pragma unroll
                    for (int wi = 0; wi < NWMAX; ++wi) {
                        if (wi + 1 < NWMAX) {
                            const double* __restrict__ rn = rb + width_of(wi + 1);
#pragma unroll
                            for (int i = 0; i < RN_PER; ++i) buf[(wi + 1) & 1][i] = rn[RN_THREADS * i];
                        }
                        __builtin_amdgcn_sched_barrier(0);
                        gate |= reaches(wi, max_of(buf[wi & 1])) << wi;
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
#pragma unroll
                    for (int wi = 0; wi < NWMAX; ++wi) {
                        if (wi >= nw) continue;  // uniform
                        const double* __restrict__ rw = rb + width_of(wi);
                        double v[RN_PER];
#pragma unroll
                        for (int i = 0; i < RN_PER; ++i) v[i] = rw[RN_THREADS * i];
                        gate |= reaches(wi, max_of(v)) << wi;
                    }
                }
                // pass 2: exact first-maximum tracking per lane, for the widths whose bit is
                // set. A lane whose sums stay below the block's largest tracked sum cannot hold
                // the block's first maximum (every earlier position holds less than it);
                // reaching it (>=) covers a tie with another wave's later position in this
                // same segment. The test runs on the upper half of the order key, which can
                // only let more lanes through.
#pragma unroll
                for (int wi = 0; wi < NWMAX; ++wi) {
                    if (!((gate >> wi) & 1u)) continue;  // uniform
                    const double* __restrict__ rw = rb + width_of(wi);
                    const double old = bestv[wi];
                    double bv = old;
                    int32_t bc = bestt[wi] - frame;
#pragma unroll
                    for (int i = 0; i < RN_PER; ++i) {
                        double di = rw[RN_THREADS * i] - q0[i];
                        if constexpr (!decltype(all_valid)::value) di = (i < nvalid) ? di : -INFINITY;
                        const bool c = di > bv;  // strict: a tie keeps the lane's earlier position
                        bv = c ? di : bv;
                        bc = c ? i : bc;
                        // keeps each position select next to its compare (otherwise the compiler
                        // hoists all compares and parks their masks in spilled registers)
                        asm volatile("" : "+v"(bv), "+v"(bc));
                    }
                    const uint32_t wk = wave_max_u32(static_cast<uint32_t>(order_key(bv) >> 32));
                    if (lane == 0) atomicMax(&tau_hi[wi], wk);
                    bestv[wi] = bv;
                    bestt[wi] = bc + frame;
                }
            };
            if (full)
                boxes(std::true_type{});
            else
                boxes(std::false_type{});
        }
    }
    // the block's first maximum per width: larger sum, then smaller position
#pragma unroll
    for (int wi = 0; wi < NWMAX; ++wi) {
        if (wi >= nw) continue;  // uniform
        const int32_t c = bestt[wi];
        const int32_t t = (c == NB_NONE) ? NB_NOPOS
                                         : static_cast<int32_t>(t0 + static_cast<int64_t>(c >> 3) * RN_SEG + tid +
                                                                RN_THREADS * (c & 7));
        const double v = wave_max(bestv[wi]);
        const int32_t tw = wave_min(bestv[wi] == v ? t : NB_NOPOS);
        if (lane == 0) {
            wv[wi][wave] = v;
            wt[wi][wave] = tw;
        }
    }
    const double s_blk = block_sum(s_acc, red);  // its barriers also order wv/wt
    if (tid < nw) {
        double v = wv[tid][0];
        int32_t t = wt[tid][0];
        for (int k = 1; k < RN_THREADS / 64; ++k) {
            const double ov = wv[tid][k];
            const int32_t ot = wt[tid][k];
            if (ov > v || (ov == v && ot < t)) {
                v = ov;
                t = ot;
            }
        }
        part_v[logical * MF_MAXW + tid] = v;
        part_t[logical * MF_MAXW + tid] = t;
    }
    if (tid == 0) part_s[logical] = s_blk;
}

// Per row (one 64-thread block): S from the chunk sums in chunk order; per width (one thread
// each) the first maximum over the chunks and conv_w; then the template-major first maximum.
__global__ __launch_bounds__(64) void k_nb_final(const NbRow* __restrict__ rows, int64_t nb_row,
                                                 const int32_t* __restrict__ widths, int nw,
                                                 const double* __restrict__ part_s,
                                                 const double* __restrict__ part_v,
                                                 const int32_t* __restrict__ part_t, float* __restrict__ snr,
                                                 int32_t* __restrict__ peak, int32_t* __restrict__ widx) {
    __shared__ double S_sh;
    __shared__ float cv_sh[MF_MAXW];
    __shared__ int32_t t_sh[MF_MAXW];
    const int64_t row = blockIdx.x;
    const NbRow r = rows[row];
    if (r.odd) return;  // the FFT path writes this row's result
    const int64_t nch = (r.n + RN_CHB - 1) / RN_CHB;
    const int lane = threadIdx.x;
    if (lane == 0) {
        double S = 0.0;
        for (int64_t c = 0; c < nch; ++c) S += part_s[row * nb_row + c];
        S_sh = S;
    }
    __syncthreads();
    if (lane < nw) {
        double bv = -INFINITY;
        int32_t bt = NB_NOPOS;
        for (int64_t c = 0; c < nch; ++c) {
            const int64_t i = (row * nb_row + c) * MF_MAXW + lane;
            const double v = part_v[i];
            if (v > bv) {  // chunks in time order: a tie keeps the earlier one
                bv = v;
                bt = part_t[i];
            }
        }
        const double w = widths[lane], ng = static_cast<double>(r.ng);
        cv_sh[lane] = static_cast<float>((bv - w * S_sh / ng) * (1.0 / sqrt(w * (1.0 - w / ng))));
        t_sh[lane] = bt;
    }
    __syncthreads();
    if (lane == 0) {
        Best best;
        best.val = -FLT_MAX;
        best.w = 0x7fffffff;
        best.t = INT64_MAX;
        for (int wi = 0; wi < nw; ++wi) {
            if (t_sh[wi] == NB_NOPOS) continue;
            if (better(cv_sh[wi], wi, t_sh[wi], best)) {
                best.val = cv_sh[wi];
                best.w = wi;
                best.t = t_sh[wi];
            }
        }
        snr[row] = best.val;
        peak[row] = static_cast<int32_t>(best.t);
        widx[row] = best.w;
    }
}

// ---------------------------------------------------------------- template bank

using spp::TempDesc;

__global__ __launch_bounds__(MF_THREADS) void k_template_conv(
    const float* __restrict__ z, int64_t n, int64_t ng, const double* __restrict__ S,
    const float* __restrict__ taps, const TempDesc* __restrict__ temps, float* __restrict__ convs) {
    __shared__ float zs[CT_B + CT_TAPS];
    __shared__ float hs[CT_TAPS];
    const int64_t a = static_cast<int64_t>(blockIdx.x) * CT_B;
    const TempDesc td = temps[blockIdx.y];
    const int tid = threadIdx.x;
    // zt[(t - ref + m) mod Ng] for t in [a, a + CT_B), m < len.
    const int nel = CT_B + td.len - 1;
    for (int k = tid; k < nel; k += MF_THREADS) {
        int64_t i = a - td.ref + k;  // in [-ref, n + CT_B + CT_TAPS)
        if (i < 0) i += ng;
        i = wrap_index(i, ng);
        zs[k] = z[wrap_index(i, n)];  // ng < 2n
    }
    for (int m = tid; m < td.len; m += MF_THREADS) hs[m] = taps[td.off + m];
    __syncthreads();
    double mu, inv;
    spp::temp_norm(td, ng, &mu, &inv);
    const double mus = mu * S[0];
    for (int u = tid; u < CT_B; u += MF_THREADS) {
        const int64_t t = a + u;
        if (t >= n) break;
        double acc = 0.0;
        for (int m = 0; m < td.len; ++m) acc += static_cast<double>(hs[m]) * static_cast<double>(zs[u + m]);
        convs[static_cast<int64_t>(blockIdx.y) * n + t] = static_cast<float>((acc - mus) * inv);
    }
}

__global__ __launch_bounds__(MF_THREADS) void k_circ_sum1(const float* __restrict__ z, int64_t n,
                                                          int64_t ng, double* __restrict__ S) {
    __shared__ double red[MF_THREADS];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < ng; i += MF_THREADS) s += static_cast<double>(z[wrap_index(i, n)]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) S[0] = s;
}

using spp::good_size_real;

struct NormLayout {
    int64_t off_rows = 0, bytes = 0;
};

NormLayout norm_layout(int64_t nrows) {
    NormLayout l;
    int64_t off = 0;
    l.off_rows = off; off = align256(off + static_cast<int64_t>(sizeof(RowDesc)) * nrows);
    l.bytes = off;
    return l;
}

}  // namespace

extern "C" int spp_running_normalise_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                                    int32_t window, int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(nrows > 0 && row_len_host != nullptr && bytes != nullptr, "invalid argument");
    SPP_REQUIRE(window >= 1, "window must be >= 1");
    *bytes = norm_layout(nrows).bytes;
    return SPP_OK;
}

extern "C" int spp_running_normalise(const float* x, int64_t ldx, int64_t nrows,
                                     const int64_t* row_len_host, int32_t window, float* z,
                                     int64_t ldz, void* workspace, int64_t workspace_bytes,
                                     void* stream) {
    spp::clear_error();
    SPP_REQUIRE(x != nullptr && z != nullptr && workspace != nullptr && row_len_host != nullptr,
                "null argument");
    SPP_REQUIRE(nrows > 0, "nrows must be > 0");
    SPP_REQUIRE(window >= 1, "window must be >= 1");
    SPP_REQUIRE(static_cast<const void*>(x) != static_cast<const void*>(z), "z must not alias x");
    std::vector<RowDesc> rows(nrows);
    int64_t maxn = 0;
    for (int64_t i = 0; i < nrows; ++i) {
        SPP_REQUIRE(row_len_host[i] >= 1 && row_len_host[i] <= ldx && row_len_host[i] <= ldz,
                    "row %lld: length %lld outside [1, ld]", static_cast<long long>(i),
                    static_cast<long long>(row_len_host[i]));
        rows[i].n = row_len_host[i];
        maxn = std::max(maxn, rows[i].n);
    }
    const NormLayout l = norm_layout(nrows);
    SPP_REQUIRE(workspace_bytes >= l.bytes, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(l.bytes));
    const int64_t nb_row = spp::ceil_div(maxn, int64_t(RN_CHB));
    const int64_t nblocks = nrows * nb_row;
    SPP_REQUIRE(nblocks + 8 <= 0x7fffffffLL, "too many rows x samples for one launch");
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    RowDesc* drows = reinterpret_cast<RowDesc*>(ws + l.off_rows);
    SPP_HIP(hipMemcpyAsync(drows, rows.data(), sizeof(RowDesc) * nrows, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_norm, dim3(static_cast<unsigned>(spp::ceil_div(nblocks, int64_t(8)) * 8)),
                       dim3(RN_THREADS), 0, st, x, ldx, drows, nb_row, nblocks, window, z, ldz);
    SPP_LAUNCHED("k_norm");
    // The row descriptors are read by the kernels above: keep the host copy alive.
    SPP_HIP(hipStreamSynchronize(st));
    return SPP_OK;
}

namespace {
struct MfLayout {
    int64_t chunk_pitch = 0, off_rows = 0, off_S = 0, off_w = 0, off_part = 0;
    // rows whose good size is odd (see spp_fft.hpp): descriptors, taps, convs, FFT space
    int64_t odd_n = 0, off_td = 0, off_taps = 0, off_convs = 0, off_fft = 0, fft_bytes = 0;
    int64_t bytes = 0;
};
MfLayout mf_layout(int64_t nrows, const int64_t* row_len, const int32_t* widths, int32_t nw) {
    MfLayout l;
    int64_t max_len = 1, sumw = 0;
    for (int64_t i = 0; i < nrows; ++i) {
        max_len = std::max(max_len, row_len[i]);
        const int64_t ng = good_size_real(row_len[i]);
        if ((ng & 1) && ng > row_len[i]) {
            l.odd_n = std::max(l.odd_n, row_len[i]);
            l.fft_bytes = std::max(l.fft_bytes, spp::odd_conv_workspace(row_len[i], nw));
        }
    }
    for (int i = 0; i < nw; ++i) sumw += widths[i];
    l.chunk_pitch = std::max<int64_t>(1, (max_len + MF_B - 1) / MF_B);
    int64_t off = 0;
    l.off_rows = off; off = align256(off + static_cast<int64_t>(sizeof(MfRow)) * nrows);
    l.off_S = off; off = align256(off + 8 * MF_NSP * nrows);
    l.off_w = off; off = align256(off + 4 * MF_MAXW);
    l.off_part = off; off = align256(off + static_cast<int64_t>(sizeof(Best)) * nrows * l.chunk_pitch);
    if (l.odd_n > 0) {
        l.off_td = off; off = align256(off + static_cast<int64_t>(sizeof(TempDesc)) * nw);
        l.off_taps = off; off = align256(off + 4 * sumw);
        l.off_convs = off; off = align256(off + 4 * l.odd_n * nw);
        l.off_fft = off; off = align256(off + l.fft_bytes);
    }
    l.bytes = off;
    return l;
}

// The first maximum over convs[w][t] (t in one MF_B chunk of a row of n samples) in
// template-major order, as k_boxcar_part's partial for that chunk.
__global__ __launch_bounds__(MF_THREADS) void k_conv_argmax_part(const float* __restrict__ convs,
                                                                 int64_t ldc, int64_t n, int nw,
                                                                 Best* __restrict__ part) {
    __shared__ Best bred[MF_THREADS];
    const int64_t a = static_cast<int64_t>(blockIdx.x) * MF_B;
    const int tid = threadIdx.x;
    Best best;
    best.val = -FLT_MAX;
    best.w = 0x7fffffff;
    best.t = INT64_MAX;
    for (int wi = 0; wi < nw; ++wi) {
        for (int u = tid; u < MF_B; u += MF_THREADS) {
            const int64_t t = a + u;
            if (t >= n) break;
            const float v = convs[static_cast<int64_t>(wi) * ldc + t];
            if (better(v, wi, t, best)) {
                best.val = v;
                best.w = wi;
                best.t = t;
            }
        }
    }
    bred[tid] = best;
    __syncthreads();
    for (int s = MF_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const Best o = bred[tid + s];
            if (better(o.val, o.w, o.t, bred[tid])) bred[tid] = o;
        }
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = bred[0];
}
}  // namespace

extern "C" int spp_boxcar_mf_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                            const int32_t* widths_host, int32_t nwidths, int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(nrows > 0 && row_len_host != nullptr && widths_host != nullptr && bytes != nullptr,
                "invalid argument");
    SPP_REQUIRE(nwidths >= 1 && nwidths <= MF_MAXW, "nwidths must be in [1, %d]", MF_MAXW);
    for (int64_t i = 0; i < nrows; ++i) SPP_REQUIRE(row_len_host[i] >= 1, "row lengths must be >= 1");
    for (int i = 0; i < nwidths; ++i)
        SPP_REQUIRE(widths_host[i] >= 1 && widths_host[i] <= MF_WMAX, "width %d outside [1, %d]",
                    widths_host[i], MF_WMAX);
    *bytes = mf_layout(nrows, row_len_host, widths_host, nwidths).bytes;
    return SPP_OK;
}

extern "C" int spp_boxcar_mf(const float* z, int64_t ldz, int64_t nrows, const int64_t* row_len_host,
                             const int32_t* widths_host, int32_t nwidths, float* snr, int32_t* peak,
                             int32_t* widx, void* workspace, int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(z != nullptr && snr != nullptr && peak != nullptr && widx != nullptr &&
                    workspace != nullptr && widths_host != nullptr && row_len_host != nullptr,
                "null argument");
    SPP_REQUIRE(nrows > 0 && nrows <= 65535, "nrows must be in [1, 65535]");
    SPP_REQUIRE(nwidths >= 1 && nwidths <= MF_MAXW, "nwidths must be in [1, %d]", MF_MAXW);
    int wmax = 0;
    for (int i = 0; i < nwidths; ++i) {
        SPP_REQUIRE(widths_host[i] >= 1 && widths_host[i] <= MF_WMAX, "width %d outside [1, %d]",
                    widths_host[i], MF_WMAX);
        wmax = std::max(wmax, widths_host[i]);
    }
    std::vector<MfRow> rows(nrows);
    std::vector<int64_t> odd_rows;
    for (int64_t i = 0; i < nrows; ++i) {
        const int64_t n = row_len_host[i];
        SPP_REQUIRE(n >= 1 && n <= ldz, "row %lld: length %lld outside [1, ldz]",
                    static_cast<long long>(i), static_cast<long long>(n));
        SPP_REQUIRE(wmax <= n, "row %lld: width %d larger than the row (%lld)",
                    static_cast<long long>(i), wmax, static_cast<long long>(n));
        rows[i].n = n;
        rows[i].ng = good_size_real(n);
        // The reference's default-length irfft gives Ng - 1 samples for an odd Ng, too few
        // to fill a row of Ng samples (core/kernels.py:904-905): it raises there.
        SPP_REQUIRE(!((rows[i].ng & 1) && rows[i].ng == n),
                    "row %lld: length %lld is an odd FFT good size; the reference's irfft returns "
                    "%lld samples and cannot fill the row",
                    static_cast<long long>(i), static_cast<long long>(n), static_cast<long long>(n - 1));
        if (rows[i].ng & 1) odd_rows.push_back(i);
    }
    const MfLayout l = mf_layout(nrows, row_len_host, widths_host, nwidths);
    SPP_REQUIRE(workspace_bytes >= l.bytes, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(l.bytes));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    MfRow* drows = reinterpret_cast<MfRow*>(ws + l.off_rows);
    double* S = reinterpret_cast<double*>(ws + l.off_S);
    int32_t* dw = reinterpret_cast<int32_t*>(ws + l.off_w);
    Best* part = reinterpret_cast<Best*>(ws + l.off_part);
    SPP_HIP(hipMemcpyAsync(drows, rows.data(), sizeof(MfRow) * nrows, hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(dw, widths_host, 4 * nwidths, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_circ_sum, dim3(MF_NSP, static_cast<unsigned>(nrows)), dim3(MF_THREADS), 0, st, z,
                       ldz, drows, S);
    SPP_LAUNCHED("k_circ_sum");
    hipLaunchKernelGGL(k_boxcar_part,
                       dim3(static_cast<unsigned>(l.chunk_pitch), static_cast<unsigned>(nrows)),
                       dim3(MF_THREADS), 0, st, z, ldz, drows, S, dw, nwidths, wmax, l.chunk_pitch, part);
    SPP_LAUNCHED("k_boxcar_part");
    // Rows with an odd good size: the template bank through the FFT path, then their chunk
    // partials (k_boxcar_part left those rows' partials neutral).
    std::vector<TempDesc> td(nwidths);
    std::vector<float> taps;
    if (!odd_rows.empty()) {
        for (int i = 0; i < nwidths; ++i) {
            const int32_t w = widths_host[i];
            td[i].len = w;
            td[i].ref = 0;
            td[i].off = static_cast<int64_t>(taps.size());
            td[i].sh = w;
            td[i].sh2 = w;
            taps.insert(taps.end(), w, 1.0f);
        }
        TempDesc* dtd = reinterpret_cast<TempDesc*>(ws + l.off_td);
        float* dtaps = reinterpret_cast<float*>(ws + l.off_taps);
        float* convs = reinterpret_cast<float*>(ws + l.off_convs);
        SPP_HIP(hipMemcpyAsync(dtd, td.data(), sizeof(TempDesc) * nwidths, hipMemcpyHostToDevice, st));
        SPP_HIP(hipMemcpyAsync(dtaps, taps.data(), 4 * taps.size(), hipMemcpyHostToDevice, st));
        for (const int64_t row : odd_rows) {
            const int64_t n = rows[row].n;
            const int rc = spp::odd_convolve(z + row * ldz, n, dtaps, dtd, nwidths, convs, l.odd_n,
                                             ws + l.off_fft, l.fft_bytes, st);
            if (rc != SPP_OK) return rc;
            hipLaunchKernelGGL(k_conv_argmax_part, dim3(static_cast<unsigned>(spp::ceil_div(n, MF_B))),
                               dim3(MF_THREADS), 0, st, convs, l.odd_n, n, nwidths, part + row * l.chunk_pitch);
            SPP_LAUNCHED("k_conv_argmax_part");
        }
    }
    hipLaunchKernelGGL(k_boxcar_final, dim3(static_cast<unsigned>(spp::ceil_div(nrows, 64))), dim3(64), 0,
                       st, part, nrows, l.chunk_pitch, drows, snr, peak, widx);
    SPP_LAUNCHED("k_boxcar_final");
    SPP_HIP(hipStreamSynchronize(st));
    return SPP_OK;
}

namespace {
// Workspace of spp_normalise_boxcar_mf: row descriptors, widths and the chunk partials of
// k_norm_boxcar; for rows with an odd good size one row's z, taps, convs, FFT space and
// chunk partials, reused row after row.
struct NbLayout {
    int64_t nb_row = 0, off_rows = 0, off_w = 0, off_s = 0, off_v = 0, off_t = 0;
    int64_t odd_n = 0, nodd = 0, fft_bytes = 0;
    int64_t off_rd = 0, off_mrow = 0, off_z = 0, off_td = 0, off_taps = 0, off_convs = 0, off_fft = 0,
            off_part = 0;
    int64_t bytes = 0;
};
NbLayout nb_layout(int64_t nrows, const int64_t* row_len, const int32_t* widths, int32_t nw) {
    NbLayout l;
    int64_t max_len = 1, sumw = 0;
    for (int64_t i = 0; i < nrows; ++i) {
        max_len = std::max(max_len, row_len[i]);
        const int64_t ng = good_size_real(row_len[i]);
        if ((ng & 1) && ng > row_len[i]) {
            ++l.nodd;
            l.odd_n = std::max(l.odd_n, row_len[i]);
            l.fft_bytes = std::max(l.fft_bytes, spp::odd_conv_workspace(row_len[i], nw));
        }
    }
    for (int i = 0; i < nw; ++i) sumw += widths[i];
    l.nb_row = std::max<int64_t>(1, spp::ceil_div(max_len, RN_CHB));
    int64_t off = 0;
    l.off_rows = off; off = align256(off + static_cast<int64_t>(sizeof(NbRow)) * nrows);
    l.off_w = off; off = align256(off + 4 * MF_MAXW);
    l.off_s = off; off = align256(off + 8 * nrows * l.nb_row);
    l.off_v = off; off = align256(off + 8 * nrows * l.nb_row * MF_MAXW);
    l.off_t = off; off = align256(off + 4 * nrows * l.nb_row * MF_MAXW);
    if (l.nodd > 0) {
        l.off_rd = off; off = align256(off + static_cast<int64_t>(sizeof(RowDesc)) * l.nodd);
        l.off_mrow = off; off = align256(off + static_cast<int64_t>(sizeof(MfRow)) * l.nodd);
        l.off_z = off; off = align256(off + 4 * l.odd_n);
        l.off_td = off; off = align256(off + static_cast<int64_t>(sizeof(TempDesc)) * nw);
        l.off_taps = off; off = align256(off + 4 * sumw);
        l.off_convs = off; off = align256(off + 4 * l.odd_n * nw);
        l.off_fft = off; off = align256(off + l.fft_bytes);
        l.off_part = off; off = align256(off + static_cast<int64_t>(sizeof(Best)) * spp::ceil_div(l.odd_n, MF_B));
    }
    l.bytes = off;
    return l;
}

template <int NWMAX, bool EXACT>
void launch_norm_boxcar(bool aligned, dim3 grid, hipStream_t st, const float* x, int64_t ldx, const NbRow* rows,
                        int64_t nb_row, int64_t nblocks, int32_t window, const int32_t* widths, int nw,
                        int32_t wmax, double* ps, double* pv, int32_t* pt) {
    if (aligned)
        hipLaunchKernelGGL((k_norm_boxcar<NWMAX, EXACT, true>), grid, dim3(RN_THREADS), 0, st, x, ldx, rows, nb_row,
                           nblocks, window, widths, nw, wmax, ps, pv, pt);
    else
        hipLaunchKernelGGL((k_norm_boxcar<NWMAX, EXACT, false>), grid, dim3(RN_THREADS), 0, st, x, ldx, rows,
                           nb_row, nblocks, window, widths, nw, wmax, ps, pv, pt);
}
}  // namespace

extern "C" int spp_normalise_boxcar_mf_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                                      const int32_t* widths_host, int32_t nwidths,
                                                      int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(nrows > 0 && row_len_host != nullptr && widths_host != nullptr && bytes != nullptr,
                "invalid argument");
    SPP_REQUIRE(nwidths >= 1 && nwidths <= MF_MAXW, "nwidths must be in [1, %d]", MF_MAXW);
    for (int64_t i = 0; i < nrows; ++i) SPP_REQUIRE(row_len_host[i] >= 1, "row lengths must be >= 1");
    for (int i = 0; i < nwidths; ++i)
        SPP_REQUIRE(widths_host[i] >= 1 && widths_host[i] <= MF_WMAX, "width %d outside [1, %d]",
                    widths_host[i], MF_WMAX);
    *bytes = nb_layout(nrows, row_len_host, widths_host, nwidths).bytes;
    return SPP_OK;
}

extern "C" int spp_normalise_boxcar_mf(const float* x, int64_t ldx, int64_t nrows, const int64_t* row_len_host,
                                       int32_t window, const int32_t* widths_host, int32_t nwidths, float* snr,
                                       int32_t* peak, int32_t* widx, void* workspace, int64_t workspace_bytes,
                                       void* stream) {
    spp::clear_error();
    SPP_REQUIRE(x != nullptr && snr != nullptr && peak != nullptr && widx != nullptr && workspace != nullptr &&
                    widths_host != nullptr && row_len_host != nullptr,
                "null argument");
    SPP_REQUIRE(nrows > 0, "nrows must be > 0");
    SPP_REQUIRE(window >= 1, "window must be >= 1");
    SPP_REQUIRE(nwidths >= 1 && nwidths <= MF_MAXW, "nwidths must be in [1, %d]", MF_MAXW);
    int wmax = 0;
    for (int i = 0; i < nwidths; ++i) {
        SPP_REQUIRE(widths_host[i] >= 1 && widths_host[i] <= MF_WMAX, "width %d outside [1, %d]",
                    widths_host[i], MF_WMAX);
        wmax = std::max(wmax, widths_host[i]);
    }
    std::vector<NbRow> rows(nrows);
    std::vector<int64_t> odd_rows;
    for (int64_t i = 0; i < nrows; ++i) {
        const int64_t n = row_len_host[i];
        SPP_REQUIRE(n >= 1 && n <= ldx && n <= 0x7fffffffLL, "row %lld: length %lld outside [1, min(ldx, 2^31 - 1)]",
                    static_cast<long long>(i), static_cast<long long>(n));
        SPP_REQUIRE(wmax <= n, "row %lld: width %d larger than the row (%lld)", static_cast<long long>(i), wmax,
                    static_cast<long long>(n));
        rows[i].n = n;
        rows[i].ng = good_size_real(n);
        SPP_REQUIRE(!((rows[i].ng & 1) && rows[i].ng == n),
                    "row %lld: length %lld is an odd FFT good size; the reference's irfft returns "
                    "%lld samples and cannot fill the row",
                    static_cast<long long>(i), static_cast<long long>(n), static_cast<long long>(n - 1));
        rows[i].odd = rows[i].ng & 1;
        if (rows[i].odd) odd_rows.push_back(i);
    }
    const NbLayout l = nb_layout(nrows, row_len_host, widths_host, nwidths);
    SPP_REQUIRE(workspace_bytes >= l.bytes, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(l.bytes));
    const int64_t nblocks = nrows * l.nb_row;
    SPP_REQUIRE(nblocks + 8 <= 0x7fffffffLL, "too many rows x samples for one launch");
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    NbRow* drows = reinterpret_cast<NbRow*>(ws + l.off_rows);
    int32_t* dw = reinterpret_cast<int32_t*>(ws + l.off_w);
    double* ps = reinterpret_cast<double*>(ws + l.off_s);
    double* pv = reinterpret_cast<double*>(ws + l.off_v);
    int32_t* pt = reinterpret_cast<int32_t*>(ws + l.off_t);
    // row descriptors and widths (zero-padded to MF_MAXW) in one copy: they are adjacent in
    // the workspace
    std::vector<char> hdesc(l.off_w + 4 * MF_MAXW, 0);
    std::memcpy(hdesc.data() + l.off_rows, rows.data(), sizeof(NbRow) * nrows);
    std::memcpy(hdesc.data() + l.off_w, widths_host, 4 * nwidths);
    SPP_HIP(hipMemcpyAsync(drows, hdesc.data(), hdesc.size(), hipMemcpyHostToDevice, st));
    // float4 loads of the three sample streams when every row and window offset is 16-byte aligned
    const int64_t L = window / 2;
    const bool aligned = (reinterpret_cast<uintptr_t>(x) & 15) == 0 && ldx % 4 == 0 && L % 4 == 0 &&
                         (window - L) % 4 == 0;
    const dim3 grid(static_cast<unsigned>(spp::ceil_div(nblocks, int64_t(8)) * 8));
    // the width loops are unrolled: exactly 9 widths (the doubling ladder 1..256), up to 12,
    // up to 32
    if (nwidths == 9)
        launch_norm_boxcar<9, true>(aligned, grid, st, x, ldx, drows, l.nb_row, nblocks, window, dw, nwidths, wmax,
                                    ps, pv, pt);
    else if (nwidths <= 12)
        launch_norm_boxcar<12, false>(aligned, grid, st, x, ldx, drows, l.nb_row, nblocks, window, dw, nwidths, wmax,
                                      ps, pv, pt);
    else
        launch_norm_boxcar<MF_MAXW, false>(aligned, grid, st, x, ldx, drows, l.nb_row, nblocks, window, dw, nwidths,
                                           wmax, ps, pv, pt);
    SPP_LAUNCHED("k_norm_boxcar");
    hipLaunchKernelGGL(k_nb_final, dim3(static_cast<unsigned>(nrows)), dim3(64), 0, st, drows, l.nb_row, dw,
                       nwidths, ps, pv, pt, snr, peak, widx);
    SPP_LAUNCHED("k_nb_final");
    // Rows with an odd good size, one after the other: z through k_norm, the template bank
    // through the FFT path, then the chunk partials and the row's result.
    std::vector<RowDesc> rd(l.nodd);
    std::vector<MfRow> mrow(l.nodd);
    if (!odd_rows.empty()) {
        std::vector<TempDesc> td(nwidths);
        std::vector<float> taps;
        for (int i = 0; i < nwidths; ++i) {
            const int32_t w = widths_host[i];
            td[i].len = w;
            td[i].ref = 0;
            td[i].off = static_cast<int64_t>(taps.size());
            td[i].sh = w;
            td[i].sh2 = w;
            taps.insert(taps.end(), w, 1.0f);
        }
        for (size_t k = 0; k < odd_rows.size(); ++k) {
            rd[k].n = rows[odd_rows[k]].n;
            mrow[k].n = rows[odd_rows[k]].n;
            mrow[k].ng = rows[odd_rows[k]].ng;
        }
        RowDesc* drd = reinterpret_cast<RowDesc*>(ws + l.off_rd);
        MfRow* dmrow = reinterpret_cast<MfRow*>(ws + l.off_mrow);
        float* zrow = reinterpret_cast<float*>(ws + l.off_z);
        TempDesc* dtd = reinterpret_cast<TempDesc*>(ws + l.off_td);
        float* dtaps = reinterpret_cast<float*>(ws + l.off_taps);
        float* convs = reinterpret_cast<float*>(ws + l.off_convs);
        Best* part = reinterpret_cast<Best*>(ws + l.off_part);
        SPP_HIP(hipMemcpyAsync(drd, rd.data(), sizeof(RowDesc) * l.nodd, hipMemcpyHostToDevice, st));
        SPP_HIP(hipMemcpyAsync(dmrow, mrow.data(), sizeof(MfRow) * l.nodd, hipMemcpyHostToDevice, st));
        SPP_HIP(hipMemcpyAsync(dtd, td.data(), sizeof(TempDesc) * nwidths, hipMemcpyHostToDevice, st));
        SPP_HIP(hipMemcpyAsync(dtaps, taps.data(), 4 * taps.size(), hipMemcpyHostToDevice, st));
        for (size_t k = 0; k < odd_rows.size(); ++k) {
            const int64_t row = odd_rows[k];
            const int64_t n = rows[row].n;
            const int64_t nb1 = spp::ceil_div(n, int64_t(RN_CHB));
            hipLaunchKernelGGL(k_norm, dim3(static_cast<unsigned>(spp::ceil_div(nb1, int64_t(8)) * 8)),
                               dim3(RN_THREADS), 0, st, x + row * ldx, ldx, drd + k, nb1, nb1, window, zrow, n);
            SPP_LAUNCHED("k_norm");
            const int rc = spp::odd_convolve(zrow, n, dtaps, dtd, nwidths, convs, l.odd_n, ws + l.off_fft,
                                             l.fft_bytes, st);
            if (rc != SPP_OK) return rc;
            hipLaunchKernelGGL(k_conv_argmax_part, dim3(static_cast<unsigned>(spp::ceil_div(n, MF_B))),
                               dim3(MF_THREADS), 0, st, convs, l.odd_n, n, nwidths, part);
            SPP_LAUNCHED("k_conv_argmax_part");
            hipLaunchKernelGGL(k_boxcar_final, dim3(1), dim3(64), 0, st, part, int64_t(1), int64_t(0), dmrow + k,
                               snr + row, peak + row, widx + row);
            SPP_LAUNCHED("k_boxcar_final");
        }
        // The host copies above are read by the stream's kernels: wait before they go.
        SPP_HIP(hipStreamSynchronize(st));
        return SPP_OK;
    }
    SPP_HIP(hipStreamSynchronize(st));
    return SPP_OK;
}

namespace {

// Scratch of the allocation-free matched filter (spp_boxcar_mf_noalloc): a file-scope device
// array, one per GPU, held until the process ends. Host code never passes the array's name
// to a kernel (on the host it is only a handle); it uses the device address that
// hipGetSymbolAddress returns on each device, looked up once per device.
constexpr int64_t MF_SCRATCH_BYTES = int64_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[MF_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device id

int scratch_address(void** out) {
    int dev = 0;
    SPP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev >= static_cast<int>(g_scratch_ptr.size())) g_scratch_ptr.resize(dev + 1, nullptr);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        SPP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SPP_REQUIRE(p != nullptr, "no device address for the matched-filter scratch on device %d", dev);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SPP_OK;
}

}  // namespace

extern "C" int spp_boxcar_mf_noalloc(const float* z, int64_t ldz, int64_t nrows, const int64_t* row_len_host,
                                     const int32_t* widths_host, int32_t nwidths, float* snr, int32_t* peak,
                                     int32_t* widx, void* stream) {
    spp::clear_error();
    int64_t need = 0;
    int rc = spp_boxcar_mf_workspace_size(nrows, row_len_host, widths_host, nwidths, &need);
    if (rc != SPP_OK) return rc;
    SPP_REQUIRE(need <= MF_SCRATCH_BYTES,
                "these rows need %lld bytes of matched-filter scratch; the allocation-free entry point "
                "holds %lld (spp_boxcar_mf takes a workspace of any size)",
                static_cast<long long>(need), static_cast<long long>(MF_SCRATCH_BYTES));
    void* ws = nullptr;
    rc = scratch_address(&ws);
    if (rc != SPP_OK) return rc;
    return spp_boxcar_mf(z, ldz, nrows, row_len_host, widths_host, nwidths, snr, peak, widx, ws, MF_SCRATCH_BYTES,
                         stream);
}

namespace {
struct CtLayout {
    int64_t off_td = 0, off_taps = 0, off_fft = 0, fft_bytes = 0, bytes = 0;
};
CtLayout ct_layout(int64_t n, int32_t ntemps, int64_t total_taps) {
    CtLayout l;
    int64_t off = align256(8);
    l.off_td = off; off = align256(off + static_cast<int64_t>(sizeof(TempDesc)) * ntemps);
    l.off_taps = off; off = align256(off + 4 * total_taps);
    const int64_t ng = good_size_real(n);
    if ((ng & 1) && ng > n) {
        l.fft_bytes = spp::odd_conv_workspace(n, ntemps);
        l.off_fft = off; off = align256(off + l.fft_bytes);
    }
    l.bytes = off;
    return l;
}
}  // namespace

extern "C" int spp_convolve_templates_workspace_size(int64_t n, int32_t ntemps, int64_t total_taps,
                                                     int64_t* bytes) {
    spp::clear_error();
    SPP_REQUIRE(n > 0 && ntemps > 0 && total_taps > 0 && bytes != nullptr, "invalid argument");
    *bytes = ct_layout(n, ntemps, total_taps).bytes;
    return SPP_OK;
}

extern "C" int spp_convolve_templates(const float* data, int64_t n, const float* temps_host,
                                      const int32_t* temp_len_host, const int32_t* ref_bin_host,
                                      int32_t ntemps, float* convs, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(data != nullptr && convs != nullptr && workspace != nullptr && temps_host != nullptr &&
                    temp_len_host != nullptr && ref_bin_host != nullptr,
                "null argument");
    SPP_REQUIRE(n > 0, "n must be > 0");
    SPP_REQUIRE(ntemps > 0 && ntemps <= 65535, "ntemps must be in [1, 65535]");
    const int64_t ng = good_size_real(n);
    SPP_REQUIRE(!((ng & 1) && ng == n),
                "length %lld is an odd FFT good size: the reference's irfft returns %lld samples, "
                "too few for the output (core/kernels.py:904-905)",
                static_cast<long long>(n), static_cast<long long>(n - 1));
    std::vector<TempDesc> td(ntemps);
    int64_t off = 0;
    for (int i = 0; i < ntemps; ++i) {
        const int32_t len = temp_len_host[i];
        SPP_REQUIRE(len >= 1 && len <= CT_TAPS, "template %d: length %d outside [1, %d]", i, len, CT_TAPS);
        SPP_REQUIRE(len <= ng, "template %d longer than the padded data", i);
        SPP_REQUIRE(ref_bin_host[i] >= 0 && ref_bin_host[i] < len, "template %d: ref bin out of range", i);
        double s = 0.0, s2 = 0.0;
        for (int m = 0; m < len; ++m) {
            const double h = temps_host[off + m];
            s += h;
            s2 += h * h;
        }
        td[i].len = len;
        td[i].ref = ref_bin_host[i];
        td[i].off = off;
        td[i].sh = s;
        td[i].sh2 = s2;
        off += len;
    }
    const CtLayout l = ct_layout(n, ntemps, off);
    SPP_REQUIRE(workspace_bytes >= l.bytes, "workspace too small: %lld < %lld",
                static_cast<long long>(workspace_bytes), static_cast<long long>(l.bytes));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    double* S = reinterpret_cast<double*>(ws);
    TempDesc* dtd = reinterpret_cast<TempDesc*>(ws + l.off_td);
    float* taps = reinterpret_cast<float*>(ws + l.off_taps);
    SPP_HIP(hipMemcpyAsync(dtd, td.data(), sizeof(TempDesc) * ntemps, hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(taps, temps_host, 4 * off, hipMemcpyHostToDevice, st));
    if (l.fft_bytes > 0) {
        // Odd good size: the reference's default-length irfft (spp_fft.hpp).
        const int rc = spp::odd_convolve(data, n, taps, dtd, ntemps, convs, n, ws + l.off_fft, l.fft_bytes, st);
        if (rc != SPP_OK) return rc;
    } else {
        hipLaunchKernelGGL(k_circ_sum1, dim3(1), dim3(MF_THREADS), 0, st, data, n, ng, S);
        SPP_LAUNCHED("k_circ_sum1");
        hipLaunchKernelGGL(k_template_conv,
                           dim3(static_cast<unsigned>(spp::ceil_div(n, CT_B)), static_cast<unsigned>(ntemps)),
                           dim3(MF_THREADS), 0, st, data, n, ng, S, taps, dtd, convs);
        SPP_LAUNCHED("k_template_conv");
    }
    SPP_HIP(hipStreamSynchronize(st));
    return SPP_OK;
}
