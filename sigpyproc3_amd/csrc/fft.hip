// Float64 FFTs for the template-bank convolution at odd good sizes (spp_fft.hpp).
//
// Power-of-two transforms are Stockham passes of radix 16 (a first pass of radix 2, 4 or
// 8 takes the remaining bits): pass p reads R elements P/R apart, twiddles them by
// exp(-2 pi i r k / (Ns R)) with k = j mod Ns, takes an in-register DFT of size R and writes
// them Ns apart at (j - k) R + k; Ns is the product of the earlier radices. Twiddles come
// from one table exp(-2 pi i m / Ptw) (Ptw a power of two >= P).
//
// A DFT of any length L (first K bins) is a Bluestein chirp-z transform over a
// power-of-two length P >= L + K - 1: with w_m = exp(i pi m^2 / L),
//   X_k = conj(w_k) * sum_j (x_j conj(w_j)) w_{k-j},
// a cyclic convolution of length P. The chirp phase uses m^2 mod 2L in exact integers.
// Inverse transforms run forward on conjugated data: IFFT(Z) = conj(FFT(conj(Z))) / P.
#include "spp_fft.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int FT_THREADS = 256;
constexpr int OC_B = 1024;      // circular-sum outputs per fill block
constexpr int OC_TAPS = 2048;   // largest template length
// Device memory the two work buffers of one pass may take (templates per pass are cut
// to fit).
constexpr int64_t OC_BUDGET = int64_t(1) << 30;

// exp(-2 pi i m / 16) for m < 8: (cos, sin) of 2 pi m / 16.
constexpr double kC16[8] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                            0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674};
constexpr double kS16[8] = {0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674,
                            1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

constexpr int bitrev(int i, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

// b * exp(-2 pi i q / 16); q is a compile-time constant after unrolling.
__device__ __forceinline__ double2 rot16(double2 b, int q) {
    if (q == 0) return b;
    if (q == 4) return make_double2(b.y, -b.x);
    return cmul(b, make_double2(kC16[q], -kS16[q]));
}

// In-register forward DFT of size R (radix-2 decimation in time, natural order out).
template <int R>
__device__ __forceinline__ void dft(double2 (&v)[R]) {
    constexpr int LG = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : 4;
    double2 u[R];
#pragma unroll
    for (int i = 0; i < R; ++i) u[bitrev(i, LG)] = v[i];
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; ++k) {
                const double2 a = u[i + k];
                const double2 b = rot16(u[i + k + len / 2], k * (16 / len));
                u[i + k] = make_double2(a.x + b.x, a.y + b.y);
                u[i + k + len / 2] = make_double2(a.x - b.x, a.y - b.y);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = u[i];
}

template <int R>
__global__ __launch_bounds__(FT_THREADS) void k_fft_pass(const double2* __restrict__ in,
                                                         double2* __restrict__ out, int64_t P,
                                                         int64_t Ns, int64_t tw_step,
                                                         const double2* __restrict__ tw,
                                                         int64_t pitch) {
    const int64_t nq = P / R;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (j >= nq) return;
    const double2* __restrict__ src = in + static_cast<int64_t>(blockIdx.y) * pitch;
    double2* __restrict__ dst = out + static_cast<int64_t>(blockIdx.y) * pitch;
    const int64_t k = j & (Ns - 1);
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * nq];
    if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * k * tw_step]);
    }
    dft<R>(v);
    const int64_t d = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[d + r * Ns] = v[r];
}

__global__ void k_twiddles(double2* __restrict__ tw, int64_t Ptw) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (m >= Ptw) return;
    double s, c;
    sincospi(2.0 * static_cast<double>(m) / static_cast<double>(Ptw), &s, &c);
    tw[m] = make_double2(c, -s);
}

// exp(i pi m^2 / L), the phase reduced exactly modulo 2 pi.
__device__ __forceinline__ double2 chirp(int64_t m, int64_t L) {
    if (m < 0) m = -m;
    const int64_t r = (m * m) % (2 * L);
    double s, c;
    sincospi(static_cast<double>(r) / static_cast<double>(L), &s, &c);
    return make_double2(c, s);
}

// Bluestein kernel b: b[m mod P] = w_m for m in (-L, K), 0 elsewhere.
__global__ void k_chirp_kernel(double2* __restrict__ b, int64_t L, int64_t K, int64_t P) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (i >= P) return;
    double2 v = make_double2(0.0, 0.0);
    if (i < K) v = chirp(i, L);
    else if (i >= P - (L - 1)) v = chirp(i - P, L);
    b[i] = v;
}

// S = sum_{i<Ng} z[i mod n] in a fixed order (one block).
__global__ __launch_bounds__(FT_THREADS) void k_pad_sum(const float* __restrict__ z, int64_t n,
                                                        int64_t ng, double* __restrict__ S) {
    __shared__ double red[FT_THREADS / 64];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < ng; i += FT_THREADS) s += static_cast<double>(z[i < n ? i : i - n]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < FT_THREADS / 64; ++w) r += red[w];
        S[0] = r;
    }
}

// Row b of A (template t0 + b): the Ng-circular convolution
//   c[t] = (sum_m h_m zt[(t - ref + m) mod Ng] - mu S) * inv,  t < Ng,
// times conj(w_t) for the length-Ng chirp (zero for Ng <= t < P).
__global__ __launch_bounds__(FT_THREADS) void k_odd_fill(const float* __restrict__ z, int64_t n,
                                                         int64_t ng, const double* __restrict__ S,
                                                         const float* __restrict__ taps,
                                                         const spp::TempDesc* __restrict__ temps,
                                                         double2* __restrict__ A, int64_t P,
                                                         int64_t pitch) {
    __shared__ float zs[OC_B + OC_TAPS];
    __shared__ float hs[OC_TAPS];
    const int64_t a = static_cast<int64_t>(blockIdx.x) * OC_B;
    double2* __restrict__ Ar = A + static_cast<int64_t>(blockIdx.y) * pitch;
    const int tid = threadIdx.x;
    if (a >= ng) {
        for (int u = tid; u < OC_B; u += FT_THREADS)
            if (a + u < P) Ar[a + u] = make_double2(0.0, 0.0);
        return;
    }
    const spp::TempDesc td = temps[blockIdx.y];
    const int nel = OC_B + td.len - 1;
    for (int k = tid; k < nel; k += FT_THREADS) {
        int64_t i = (a - td.ref + k) % ng;
        if (i < 0) i += ng;
        zs[k] = z[i < n ? i : i - n];  // ng < 2n
    }
    for (int m = tid; m < td.len; m += FT_THREADS) hs[m] = taps[td.off + m];
    __syncthreads();
    double mu, inv;
    spp::temp_norm(td, ng, &mu, &inv);
    const double mus = mu * S[0];
    for (int u = tid; u < OC_B; u += FT_THREADS) {
        const int64_t t = a + u;
        if (t >= P) break;
        double2 v = make_double2(0.0, 0.0);
        if (t < ng) {
            double acc = 0.0;
            for (int m = 0; m < td.len; ++m) acc += static_cast<double>(hs[m]) * static_cast<double>(zs[u + m]);
            const double c = (acc - mus) * inv;
            const double2 w = chirp(t, ng);
            v = make_double2(c * w.x, -c * w.y);
        }
        Ar[t] = v;
    }
}

// A[b][i] = conj(A[b][i] * B[i]): the product of two spectra, conjugated so that a forward
// FFT performs the inverse transform.
__global__ void k_pmul_conj(double2* __restrict__ A, const double2* __restrict__ B, int64_t P,
                            int64_t pitch) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (i >= P) return;
    double2* __restrict__ Ar = A + static_cast<int64_t>(blockIdx.y) * pitch;
    Ar[i] = conjd(cmul(Ar[i], B[i]));
}

// From C = FFT(conj(spectrum product)) of the length-Ng Bluestein: X_k = conj(w_k C_k) / P1
// for k <= Lh, then the truncated Hermitian spectrum of length M = Ng - 1
//   Y_k = X_k (k <= Lh; imaginary parts of Y_0 and Y_Lh dropped), Y_{M-k} = conj(Y_k),
// written as the next Bluestein's input conj(Y_k w'_k) (w' the length-M chirp), zero for
// M <= k < P2.
__global__ void k_odd_mid(const double2* __restrict__ C, int64_t pitch_c, double2* __restrict__ A2,
                          int64_t pitch_a, int64_t ng, int64_t M, int64_t P1, int64_t P2) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (k >= P2) return;
    double2* __restrict__ Ar = A2 + static_cast<int64_t>(blockIdx.y) * pitch_a;
    if (k >= M) {
        Ar[k] = make_double2(0.0, 0.0);
        return;
    }
    const double2* __restrict__ Cr = C + static_cast<int64_t>(blockIdx.y) * pitch_c;
    const int64_t Lh = M / 2;
    const int64_t kk = (k <= Lh) ? k : M - k;
    double2 X = conjd(cmul(chirp(kk, ng), Cr[kk]));
    const double s = 1.0 / static_cast<double>(P1);
    X.x *= s;
    X.y *= s;
    if (kk == 0 || kk == Lh) X.y = 0.0;
    const double2 Y = (k <= Lh) ? X : conjd(X);
    Ar[k] = conjd(cmul(Y, chirp(k, M)));
}

// y_t = Re(w'_t C2_t) / (P2 M) for t < n: the inverse length-M transform's real output.
__global__ void k_odd_final(const double2* __restrict__ C2, int64_t pitch, float* __restrict__ convs,
                            int64_t ldc, int64_t n, int64_t M, int64_t P2) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * FT_THREADS + threadIdx.x;
    if (t >= n) return;
    const double2 c = C2[static_cast<int64_t>(blockIdx.y) * pitch + t];
    const double2 w = chirp(t, M);
    const double y = (w.x * c.x - w.y * c.y) / (static_cast<double>(P2) * static_cast<double>(M));
    convs[static_cast<int64_t>(blockIdx.y) * ldc + t] = static_cast<float>(y);
}

int64_t pow2_at_least(int64_t x) {
    int64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

int log2_exact(int64_t p) {
    int l = 0;
    while ((int64_t(1) << l) < p) ++l;
    return l;
}

struct OddSizes {
    int64_t ng, M, Lh, P1, P2, Pmax;
};

OddSizes odd_sizes(int64_t n) {
    OddSizes s;
    s.ng = spp::good_size_real(n);
    s.M = s.ng - 1;
    s.Lh = s.M / 2;
    s.P1 = pow2_at_least(s.ng + s.Lh);      // length ng, bins 0..Lh
    s.P2 = pow2_at_least(s.M + n - 1);      // length M, outputs 0..n-1
    s.Pmax = std::max(s.P1, s.P2);
    return s;
}

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

// Forward FFT of `batch` rows (pitch elements apart) of length P in `buf`, using `tmp`
// (same shape) as the other half of each pass. *result is the buffer holding the output.
int fft_forward(double2* buf, double2* tmp, int64_t P, int batch, int64_t pitch, const double2* tw,
                int64_t Ptw, hipStream_t st, double2** result) {
    const int lg = log2_exact(P);
    double2* cur = buf;
    double2* oth = tmp;
    int64_t Ns = 1;
    auto pass = [&](int R) -> int {
        const int64_t nq = P / R;
        const dim3 grid(static_cast<unsigned>(spp::ceil_div(nq, FT_THREADS)), static_cast<unsigned>(batch));
        const int64_t step = Ptw / (Ns * R);
        switch (R) {
            case 2: hipLaunchKernelGGL(k_fft_pass<2>, grid, dim3(FT_THREADS), 0, st, cur, oth, P, Ns, step, tw, pitch); break;
            case 4: hipLaunchKernelGGL(k_fft_pass<4>, grid, dim3(FT_THREADS), 0, st, cur, oth, P, Ns, step, tw, pitch); break;
            case 8: hipLaunchKernelGGL(k_fft_pass<8>, grid, dim3(FT_THREADS), 0, st, cur, oth, P, Ns, step, tw, pitch); break;
            default: hipLaunchKernelGGL(k_fft_pass<16>, grid, dim3(FT_THREADS), 0, st, cur, oth, P, Ns, step, tw, pitch); break;
        }
        SPP_LAUNCHED("k_fft_pass");
        std::swap(cur, oth);
        Ns *= R;
        return SPP_OK;
    };
    if (lg % 4 != 0) {
        const int rc = pass(1 << (lg % 4));
        if (rc != SPP_OK) return rc;
    }
    for (int p = 0; p < lg / 4; ++p) {
        const int rc = pass(16);
        if (rc != SPP_OK) return rc;
    }
    *result = cur;
    return SPP_OK;
}

}  // namespace

namespace spp {

int64_t good_size_real(int64_t n) {
    if (n <= 6) return n;
    // all 2^a 3^b 5^c up to 2^62, sorted, built once: the smallest one >= n
    static const std::vector<int64_t> smooth = [] {
        std::vector<int64_t> v;
        constexpr int64_t kMax = int64_t(1) << 62;
        for (int64_t f5 = 1; f5 <= kMax; f5 *= 5) {
            for (int64_t f35 = f5; f35 <= kMax; f35 *= 3) {
                for (int64_t x = f35; x <= kMax; x *= 2) {
                    v.push_back(x);
                    if (x > kMax / 2) break;
                }
                if (f35 > kMax / 3) break;
            }
            if (f5 > kMax / 5) break;
        }
        std::sort(v.begin(), v.end());
        return v;
    }();
    return *std::lower_bound(smooth.begin(), smooth.end(), n);
}

int32_t odd_conv_batch(int64_t n, int32_t ntemps) {
    const OddSizes s = odd_sizes(n);
    const int64_t per = 2 * 16 * s.Pmax;
    return static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(ntemps, OC_BUDGET / per)));
}

int64_t odd_conv_workspace(int64_t n, int32_t ntemps) {
    const OddSizes s = odd_sizes(n);
    const int64_t batch = odd_conv_batch(n, ntemps);
    return align256(16 * s.Pmax) + align256(16 * s.P1) + align256(16 * s.P2) +
           2 * align256(16 * s.Pmax * batch) + align256(8);
}

int odd_convolve(const float* z, int64_t n, const float* taps, const TempDesc* temps, int32_t ntemps,
                 float* convs, int64_t ldc, void* workspace, int64_t workspace_bytes, hipStream_t st) {
    const OddSizes s = odd_sizes(n);
    SPP_REQUIRE((s.ng & 1) && s.ng > n, "odd_convolve needs an odd good size above n");
    SPP_REQUIRE(workspace_bytes >= odd_conv_workspace(n, ntemps), "workspace too small for the odd-size path");
    const int32_t batch = odd_conv_batch(n, ntemps);
    char* ws = static_cast<char*>(workspace);
    double2* tw = reinterpret_cast<double2*>(ws);
    ws += align256(16 * s.Pmax);
    double2* B1 = reinterpret_cast<double2*>(ws);
    ws += align256(16 * s.P1);
    double2* B2 = reinterpret_cast<double2*>(ws);
    ws += align256(16 * s.P2);
    double2* X = reinterpret_cast<double2*>(ws);
    ws += align256(16 * s.Pmax * batch);
    double2* Y = reinterpret_cast<double2*>(ws);
    ws += align256(16 * s.Pmax * batch);
    double* S = reinterpret_cast<double*>(ws);
    const int64_t pitch = s.Pmax;

    hipLaunchKernelGGL(k_twiddles, dim3(static_cast<unsigned>(ceil_div(s.Pmax, FT_THREADS))), dim3(FT_THREADS), 0,
                       st, tw, s.Pmax);
    SPP_LAUNCHED("k_twiddles");
    // Spectra of the two chirp kernels.
    struct ChirpJob {
        double2* dst;
        int64_t L, K, P;
    } jobs[2] = {{B1, s.ng, s.Lh + 1, s.P1}, {B2, s.M, n, s.P2}};
    for (const ChirpJob& jb : jobs) {
        hipLaunchKernelGGL(k_chirp_kernel, dim3(static_cast<unsigned>(ceil_div(jb.P, FT_THREADS))), dim3(FT_THREADS),
                           0, st, jb.dst, jb.L, jb.K, jb.P);
        SPP_LAUNCHED("k_chirp_kernel");
        double2* res = nullptr;
        int rc = fft_forward(jb.dst, X, jb.P, 1, jb.P, tw, s.Pmax, st, &res);
        if (rc != SPP_OK) return rc;
        if (res != jb.dst) SPP_HIP(hipMemcpyAsync(jb.dst, res, 16 * jb.P, hipMemcpyDeviceToDevice, st));
    }
    hipLaunchKernelGGL(k_pad_sum, dim3(1), dim3(FT_THREADS), 0, st, z, n, s.ng, S);
    SPP_LAUNCHED("k_pad_sum");

    for (int32_t t0 = 0; t0 < ntemps; t0 += batch) {
        const int nb = std::min<int32_t>(batch, ntemps - t0);
        hipLaunchKernelGGL(k_odd_fill, dim3(static_cast<unsigned>(ceil_div(s.P1, OC_B)), static_cast<unsigned>(nb)),
                           dim3(FT_THREADS), 0, st, z, n, s.ng, S, taps, temps + t0, X, s.P1, pitch);
        SPP_LAUNCHED("k_odd_fill");
        double2* r = nullptr;
        int rc = fft_forward(X, Y, s.P1, nb, pitch, tw, s.Pmax, st, &r);
        if (rc != SPP_OK) return rc;
        const dim3 g1(static_cast<unsigned>(ceil_div(s.P1, FT_THREADS)), static_cast<unsigned>(nb));
        hipLaunchKernelGGL(k_pmul_conj, g1, dim3(FT_THREADS), 0, st, r, B1, s.P1, pitch);
        SPP_LAUNCHED("k_pmul_conj");
        double2* o = (r == X) ? Y : X;
        double2* c = nullptr;
        rc = fft_forward(r, o, s.P1, nb, pitch, tw, s.Pmax, st, &c);
        if (rc != SPP_OK) return rc;
        double2* a2 = (c == X) ? Y : X;
        const dim3 g2(static_cast<unsigned>(ceil_div(s.P2, FT_THREADS)), static_cast<unsigned>(nb));
        hipLaunchKernelGGL(k_odd_mid, g2, dim3(FT_THREADS), 0, st, c, pitch, a2, pitch, s.ng, s.M, s.P1, s.P2);
        SPP_LAUNCHED("k_odd_mid");
        rc = fft_forward(a2, c, s.P2, nb, pitch, tw, s.Pmax, st, &r);
        if (rc != SPP_OK) return rc;
        hipLaunchKernelGGL(k_pmul_conj, g2, dim3(FT_THREADS), 0, st, r, B2, s.P2, pitch);
        SPP_LAUNCHED("k_pmul_conj");
        o = (r == X) ? Y : X;
        rc = fft_forward(r, o, s.P2, nb, pitch, tw, s.Pmax, st, &c);
        if (rc != SPP_OK) return rc;
        const dim3 g3(static_cast<unsigned>(ceil_div(n, FT_THREADS)), static_cast<unsigned>(nb));
        hipLaunchKernelGGL(k_odd_final, g3, dim3(FT_THREADS), 0, st, c, pitch, convs + t0 * ldc, ldc, n, s.M, s.P2);
        SPP_LAUNCHED("k_odd_final");
    }
    return SPP_OK;
}

}  // namespace spp
