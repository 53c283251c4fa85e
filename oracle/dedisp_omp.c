/*
 * C/OpenMP restatement of sigpyproc's dedispersion kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the CPU checker and the CPU baseline timed by bench.py. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
 * library (libspp_hip.so) never links or calls it.
 *
 * spp_oracle_dedisperse restates kernels.dedisperse / dedisperse_par
 * (core/kernels.py:340-362, 542): the same loop nest -- parallel over output samples,
 * serial over channels, float64 accumulator, float32 store -- and numba's wraparound
 * of a negative row index (t + d_c < 0 -> t + d_c + nsamps).
 *
 * spp_oracle_subband restates kernels.subband (core/kernels.py:379-401) with the
 * output zeroed by the caller.
 *
 * spp_oracle_unpack restates kernels.unpack{1,2,4}_8_{big,little} (core/kernels.py:20-69).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int spp_oracle_max_threads(void) { return omp_get_max_threads(); }

void spp_oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

static int64_t max_abs_delay(const int32_t* delays, int64_t nchans) {
    int64_t m = 0;
    for (int64_t c = 0; c < nchans; ++c) {
        int64_t a = delays[c] < 0 ? -(int64_t)delays[c] : delays[c];
        if (a > m) m = a;
    }
    return m;
}

/* kernels.dedisperse on uint8 data: out[out_start + t] = f32(sum_c f64(in[(t+d_c)*nchans + c])) */
void spp_oracle_dedisperse(const uint8_t* in, float* out, const int32_t* delays, int64_t nchans,
                           int64_t nsamps, int64_t out_start) {
    const int64_t nout = nsamps - max_abs_delay(delays, nchans);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nout; ++t) {
        double summ = 0.0;
        for (int64_t c = 0; c < nchans; ++c) {
            int64_t row = t + delays[c];
            if (row < 0) row += nsamps;
            summ += (double)in[row * nchans + c];
        }
        out[out_start + t] = (float)summ;
    }
}

/* The same loop nest on float32 data (the second numba signature). */
void spp_oracle_dedisperse_f32(const float* in, float* out, const int32_t* delays, int64_t nchans,
                               int64_t nsamps, int64_t out_start) {
    const int64_t nout = nsamps - max_abs_delay(delays, nchans);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nout; ++t) {
        double summ = 0.0;
        for (int64_t c = 0; c < nchans; ++c) {
            int64_t row = t + delays[c];
            if (row < 0) row += nsamps;
            summ += (double)in[row * nchans + c];
        }
        out[out_start + t] = (float)summ;
    }
}

/*
 * DM sweep over a whole unpacked stream [nsamps][nchans] with non-negative delays:
 * row i of plane (leading dimension ld) is dedisperse with delays table[i], valid
 * length nsamps - max|d_i|; entries past the valid length are left untouched.
 */
void spp_oracle_sweep(const uint8_t* in, int64_t nsamps, int64_t nchans, const int32_t* table,
                      int64_t ndm, float* plane, int64_t ld) {
    for (int64_t i = 0; i < ndm; ++i)
        spp_oracle_dedisperse(in, plane + i * ld, table + i * nchans, nchans, nsamps, 0);
}

/* kernels.subband: out[t*nsubs + chan_to_sub[c]] += in[nchans*(t+d_c) + c], t < nsamps - maxdelay */
void spp_oracle_subband(const uint8_t* in, float* out, int64_t out_size, const int32_t* delays,
                        const int32_t* chan_to_sub, int64_t maxdelay, int64_t nchans, int64_t nsubs,
                        int64_t nsamps) {
    const int64_t n = nsamps - maxdelay;
    const int64_t size = nsamps * nchans;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; ++t) {
        for (int64_t c = 0; c < nchans; ++c) {
            int64_t idx = nchans * (t + delays[c]) + c;
            if (idx < 0) idx += size;
            int64_t o = nsubs * t + chan_to_sub[c];
            if (o < out_size) out[o] += (float)in[idx];
        }
    }
}

void spp_oracle_unpack(const uint8_t* in, int64_t nbytes, uint8_t* out, int nbits, int big) {
    const int nper = 8 / nbits;
    const int mask = (1 << nbits) - 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nbytes; ++i) {
        const int b = in[i];
        for (int k = 0; k < nper; ++k) {
            const int shift = big ? (8 - (k + 1) * nbits) : k * nbits;
            out[i * nper + k] = (uint8_t)((b >> shift) & mask);
        }
    }
}
