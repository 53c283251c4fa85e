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
//   z[t] = float32((x[t] - (m + x[0])) / s).
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
#include <mutex>
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

// Block-wide sum of a double in a fixed order (butterfly within each wave, then the wave
// totals in wave order). All threads get the result. blockDim.x is a multiple of 64.
__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) r += red[w];
    __syncthreads();
    return r;
}

// Inclusive scan of a double across one wave (fixed order).
__device__ __forceinline__ double wave_incl_scan(double v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive block scan of two doubles per thread, also returning the block totals (ta, tb):
// within each wave by shuffles, then the waves' totals from LDS in wave order. Two barriers;
// sa/sb hold one entry per wave.
__device__ void block_exscan2_total(double& a, double& b, double& ta, double& tb, double* sa, double* sb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double ia = wave_incl_scan(a), ib = wave_incl_scan(b);
    double xa = __shfl_up(ia, 1, 64), xb = __shfl_up(ib, 1, 64);
    if (lane == 0) {
        xa = 0.0;
        xb = 0.0;
    }
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
    double xa = __shfl_up(ia, 1, 64), xb = __shfl_up(ib, 1, 64);
    if (lane == 0) {
        xa = 0.0;
        xb = 0.0;
    }
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

// ---------------------------------------------------------------- running normalise

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
            const int64_t t = u0 + RN_PER * tid + j;
            p1[j] = a1;
            p2[j] = a2;
            if (t < tend) {
                const double dn = static_cast<double>(cn[j]) - ref, d0 = static_cast<double>(c0[j]) - ref;
                a1 += dn - d0;
                a2 += dn * dn - d0 * d0;
            }
        }
        double e1 = a1, e2 = a2, T1, T2;
        block_exscan2_total(e1, e2, T1, T2, sa, sb);
#pragma unroll
        for (int j = 0; j < RN_PER; ++j) {
            const int64_t t = u0 + RN_PER * tid + j;
            if (t < tend) {
                const double m = (S1 + e1 + p1[j]) * invw;
                double v = (S2 + e2 + p2[j]) * invw - m * m;
                v = v > 0.0 ? v : 0.0;
                double sd = sqrt(v);
                if (sd == 0.0) sd = 1.0;
                zr[t] = static_cast<float>((static_cast<double>(cx[j]) - (m + ref)) / sd);
            }
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
