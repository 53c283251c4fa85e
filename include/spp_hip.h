/*
 * libspp_hip -- C ABI of the MI355X (gfx950) filterbank dedispersion hot path.
 *
 * Conventions
 *   - Every entry point returns an int status: SPP_OK (0) on success, SPP_EINVAL for
 *     an invalid argument (the Python layer raises ValueError), SPP_EHIP for a HIP
 *     runtime failure (RuntimeError). spp_last_error() returns the thread-local
 *     message of the last failure on the calling thread.
 *   - Array arguments are DEVICE pointers unless the name ends in _host. The caller
 *     owns every buffer; the library allocates no device memory.
 *   - `stream` is a hipStream_t (NULL = the default stream). Work is enqueued and the
 *     call returns without synchronising, except where a comment says otherwise.
 *   - Reference paths are relative to the sigpyproc source tree (src/sigpyproc).
 */
#ifndef SPP_HIP_H
#define SPP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPP_OK 0
#define SPP_EINVAL 1
#define SPP_EHIP 2

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* spp_last_error(void);

/* ABI version of this library. */
int spp_version(void);

/*
 * Unpack 1/2/4-bit samples packed in bytes to one byte per sample.
 * Replaces kernels.unpack{1,2,4}_8_{big,little} (core/kernels.py:20-69) as called by
 * bits.unpack (io/bits.py:19-70, getattr at 68-69). big != 0: the first sample of a
 * byte is in its most significant bits. out holds nbytes * (8 / nbits) bytes.
 */
int spp_unpack(const uint8_t* in, int64_t nbytes, uint8_t* out, int nbits, int big,
               void* stream);

/*
 * Dedisperse one gulp to a time series, unpacking in registers.
 * Replaces kernels.dedisperse / kernels.dedisperse_par (core/kernels.py:340-362, 542)
 * as called by Filterbank.dedisperse (base.py:405-420), fused with the unpack of
 * FileReader.creadinto (io/fileio.py:284-292).
 *   in      gulp in file order [nsamps][nchans]: packed nbits in {1,2,4,8} with bit
 *           order `big`, or float32 samples for nbits == 32
 *   delays     int32 [nchans]
 *   max_delay  max_c |delays[c]| (the reference computes it inside the kernel,
 *              kernels.py:356; here the caller, which holds the host table, passes it)
 *   out        out[out_start + t] = sum_c in[t + delays[c]][c] for t < nsamps - max_delay;
 *              a negative row index t + d wraps to t + d + nsamps (numba semantics).
 * Integer data is summed exactly; float32 data is summed in float64 in channel order.
 * Rows outside [0, nsamps) after the wrap read as 0, so no delay value reads outside
 * the gulp.
 */
int spp_dedisperse(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                   const int32_t* delays, int64_t max_delay, float* out, int64_t out_start,
                   void* stream);

/*
 * Subband one gulp. Replaces kernels.subband (core/kernels.py:379-401) as called by
 * Filterbank.subband (base.py:1096-1105):
 *   out[t*nsubs + chan_to_sub[c]] += in[t + delays[c]][c]  for t < nsamps - maxdelay.
 * The entry point first zeroes out[0 : min(out_size, (nsamps - maxdelay) * nsubs)]
 * (the reference's np.empty buffer is never zeroed, base.py:1073; zeroing per gulp is
 * the kernel semantics its comment asks for). Writes at flat indices >= out_size are
 * dropped. nbits in {1,2,4,8} (packed, bit order `big`) or 32 (float32).
 */
int spp_subband(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                const int32_t* delays, const int32_t* chan_to_sub, int64_t maxdelay,
                int64_t nsubs, float* out, int64_t out_size, void* stream);

/*
 * DM-trial sweep: plane[i][t] = Filterbank.dedisperse(dm_i).data[t] (base.py:360-424)
 * for every row i of a host delay table, over a whole stream resident on the device.
 * No single reference function: it is the stack over DMs of kernels.dedisperse
 * (core/kernels.py:348-362) outputs, which do not depend on the gulp schedule for
 * non-negative delays. Rows have length row_len[i] = nsamps - max_c delays[i][c];
 * plane entries in [row_len[i], min(ld, ceil(max_len/T)*T)) are written as 0, with the
 * time tile T = 2048 for 1/2-bit data and 1024 for 4/8-bit data.
 */
typedef struct spp_sweep_plan spp_sweep_plan;

typedef struct {
    int64_t workspace_bytes; /* device workspace the caller must provide           */
    int64_t max_len;         /* longest row (the plane's minimum leading dimension) */
    int64_t min_len;         /* shortest row                                        */
    int64_t total_samples;   /* sum of row lengths (DM-trial samples)               */
    int64_t ndm_tiles;       /* DM tiles the planner formed                         */
    int32_t max_spread;      /* largest per-channel delay spread inside a DM tile   */
    int32_t mode;            /* field bits (4/8/16) + 32 if 32-bit lanes + 64*variant */
    int64_t tables_bytes;    /* leading part of the workspace holding the plan tables */
} spp_sweep_info;

/* Plan on the host from a host delay table int32 [ndm][nchans] (all delays >= 0). */
int spp_sweep_plan_create(const int32_t* table_host, int64_t ndm, int64_t nchans,
                          int64_t nsamps, int nbits, int big, spp_sweep_plan** plan);
int spp_sweep_plan_destroy(spp_sweep_plan* plan);
/* Block shapes of the stage-2 kernel (DM rows per wave x waves x channels per stage):
 * 1/2-bit data: 0 = 2x8x16, 1 = 2x8x20, 2 = 2x8x8 (6 waves per SIMD), 3 = 2x8x12 (6 waves per
 *               SIMD), 4 = 1x16x16 (8 waves per SIMD, default), 5 = 1x16x12 (8 waves per SIMD);
 * 4/8-bit data: 0 = 2x8x12 (6 waves per SIMD, default), 1 = 2x8x16, 2 = 4x4x8, 3 = 2x8x8,
 *               4 = 1x16x16, 5 = 1x16x12 (8 waves per SIMD).
 * All give identical results; spp_sweep_plan_set_variant picks one for a plan (the
 * autotuner times them). */
int spp_sweep_variant_count(void);
int spp_sweep_plan_set_variant(spp_sweep_plan* plan, int variant);
int spp_sweep_plan_info(const spp_sweep_plan* plan, spp_sweep_info* info);
/* Row lengths (int64 [ndm]) into host memory. */
int spp_sweep_plan_row_lengths(const spp_sweep_plan* plan, int64_t* out_host);
/* Copy the plan's device tables into `workspace` (synchronises the stream once). A workspace
 * of tables_bytes only serves spp_sweep_execute_g; the other entry points need
 * workspace_bytes (tables + the stage-1 copy for the plan's nsamps). */
int spp_sweep_plan_upload(spp_sweep_plan* plan, void* workspace, int64_t workspace_bytes,
                          void* stream);
/* Run the sweep on a packed time-major stream [nsamps][nchans] (nbits, bit order of the
 * plan) into plane [ndm][ld] (ld >= max_len). `workspace` must be the buffer the plan
 * was uploaded to. spp_sweep_execute = spp_sweep_build (stage 1: the channel-major
 * interleaved copy of the stream, kept in the workspace) + spp_sweep_run (stage 2: the
 * DM-tiled sums); the stages are exposed so a caller can time or reuse them. */
int spp_sweep_execute(const spp_sweep_plan* plan, const uint8_t* packed, float* plane,
                      int64_t ld, void* workspace, int64_t workspace_bytes, void* stream);
int spp_sweep_build(const spp_sweep_plan* plan, const uint8_t* packed, void* workspace,
                    int64_t workspace_bytes, void* stream);
int spp_sweep_run(const spp_sweep_plan* plan, float* plane, int64_t ld, void* workspace,
                  int64_t workspace_bytes, void* stream);
/* spp_sweep_execute on the first nsamps samples only (max delay of the table < nsamps <= the
 * plan's nsamps): rows are nsamps - max_c delays[i][c] long. One plan then serves gulps of any
 * length up to its own (a stream processed in overlapping gulps of varying size). */
int spp_sweep_execute_n(const spp_sweep_plan* plan, const uint8_t* packed, int64_t nsamps,
                        float* plane, int64_t ld, void* workspace, int64_t workspace_bytes,
                        void* stream);
/* spp_sweep_execute_n with the stage-1 copy in a caller's buffer `g` of at least
 * spp_sweep_g_bytes(plan, nsamps) bytes, pitched for nsamps: a server allocates it per gulp
 * (and frees it when the gulp completes) while the plan tables stay resident in `tables`. */
int spp_sweep_g_bytes(const spp_sweep_plan* plan, int64_t nsamps, int64_t* bytes);
int spp_sweep_execute_g(const spp_sweep_plan* plan, const uint8_t* packed, int64_t nsamps,
                        float* plane, int64_t ld, void* tables, int64_t tables_bytes,
                        uint32_t* g, int64_t g_bytes, void* stream);

/*
 * Batched sweep: replaces the per-call loop of Filterbank.dedisperse (base.py:405-420: one
 * kernels.dedisperse call per gulp and DM) for many filterbanks at once. The plans (uneven
 * nbits, nchans, nsamps and DM lists) run as one persistent kernel launch over a device work
 * queue of (instance, DM rows, time tiles) items, stage-1 copies included; every plane equals
 * spp_sweep_execute of its own plan bit for bit. When the stage-1 copies of all plans exceed
 * g_budget_bytes (0: no limit) the batch is cut into rounds whose copies fit half the budget,
 * and two pools alternate inside the same launch.
 *
 * Use: create -> spp_batch_get_info (workspace size) -> upload (plan tables, once) -> bind
 * (device streams, planes and leading dimensions; again after set_shape) -> run, or capture
 * once and replay: a bound batch reads its streams and writes its planes at the bound
 * addresses on every run, so repeated gulps copied into the same buffers replay one graph.
 * The plans must outlive the batch. spp_batch_errors reads and clears the device error flag
 * (nonzero: a queue wait gave up and the planes are incomplete).
 */
typedef struct spp_batch spp_batch;

typedef struct {
    int64_t workspace_bytes;     /* device workspace the caller must provide            */
    int64_t pool_bytes;          /* bytes of one stage-1 pool                           */
    int64_t build_items;         /* stage-1 queue items                                 */
    int64_t sweep_items;         /* stage-2 queue items                                 */
    int32_t rounds;              /* rounds the G budget needs (1: all copies resident)   */
    int32_t segments;            /* queue segments (2 per round)                        */
    int32_t shape;               /* block shape (spp_batch_set_shape)                   */
    int32_t time_tiles_per_item; /* time tiles of one stage-2 item                      */
    int32_t dm_rows_per_item;    /* DM rows of one stage-2 item                         */
    int32_t chans_per_stage;     /* channels per LDS stage                              */
} spp_batch_info;

int spp_batch_create(const spp_sweep_plan* const* plans, int32_t n, int64_t g_budget_bytes,
                     spp_batch** batch);
int spp_batch_destroy(spp_batch* batch);
/* Block shapes (DM rows per item x channels per stage): 0 = 16 x 12, 1 = 16 x 16, 2 = 8 x 8,
 * 3 = 16 x 8 (8-wave blocks), 4 = 16 x 16 (16 waves of one DM row, default); each stage-2 item
 * covers time_tiles_per_item consecutive time tiles (2048 samples for 1/2-bit data, 1024 for
 * 4/8-bit). All give identical planes. */
int spp_batch_shape_count(void);
int spp_batch_set_shape(spp_batch* batch, int shape, int32_t time_tiles_per_item);
int spp_batch_get_info(const spp_batch* batch, spp_batch_info* info);
/* Round of each plan (int32 [n]) into host memory. */
int spp_batch_rounds(const spp_batch* batch, int32_t* round_of_host);
int spp_batch_upload(spp_batch* batch, void* workspace, int64_t workspace_bytes, void* stream);
/* packed[i]: device stream of plan i ([nsamps][nchans], nbits, its bit order); planes[i]:
 * device plane [ndm][ld[i]], ld[i] >= the plan's max_len. Host arrays of n entries. */
int spp_batch_bind(spp_batch* batch, const uint8_t* const* packed, float* const* planes,
                   const int64_t* ld, void* workspace, int64_t workspace_bytes, void* stream);
int spp_batch_run(spp_batch* batch, void* workspace, int64_t workspace_bytes, void* stream);
/* Capture spp_batch_run into a HIP graph owned by the batch (on a private stream). */
int spp_batch_capture(spp_batch* batch, void* workspace, int64_t workspace_bytes);
int spp_batch_replay(spp_batch* batch, void* stream);
int spp_batch_errors(spp_batch* batch, void* workspace, int64_t workspace_bytes, void* stream,
                     int32_t* flags);

/*
 * LDS read-rate probe (no counterpart in the reference: the measured peak the sweep's LDS
 * reads are compared with in the benchmark's roofline). `blocks` blocks of 512 threads read
 * 32 KiB of LDS for iters rounds of 8 rows: pattern 0 = ds_read_b128 (16 aligned bytes per
 * lane), pattern 1 = ds_read2st64_b32 (the sweep's window reads, 8 bytes per lane). Bytes
 * read: blocks * 512 * iters * 8 * (16 or 8). out: uint32 [blocks * 512] (device).
 */
int spp_probe_lds(int pattern, int64_t iters, int32_t blocks, uint32_t* out, void* stream);

/*
 * Gulp pre-steps (same reads and unpack as spp_dedisperse; `in` is a gulp in file order,
 * packed nbits 1/2/4/8 with bit order `big`, or float32 for nbits == 32).
 *
 * spp_bandpass: per-channel sums of one gulp added into out[nchans] (device, float32).
 * parallel == 0 replaces kernels.extract_bpass (core/kernels.py:277-291): float32 adds in
 * time order per channel. parallel != 0 replaces kernels.extract_bpass_par
 * (core/kernels.py:294-310): out[c] = float32(float64(out[c]) + float64 sum over the gulp),
 * numba's float32 += float64. Called by Filterbank.bandpass (base.py:308-358). Integer data
 * need a device workspace of spp_bandpass_workspace_size bytes (none for float32); a
 * workspace sized for nsamps serves every shorter gulp.
 */
int spp_bandpass_workspace_size(int nbits, int64_t nchans, int64_t nsamps, int64_t* bytes);
int spp_bandpass(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps, float* out,
                 int parallel, void* workspace, int64_t workspace_bytes, void* stream);
/*
 * spp_collapse: out[out_start + t] = float32(sum over c of x[t][c]), the float64 channel sum
 * of kernels.extract_tim (core/kernels.py:255-273) as called by Filterbank.collapse
 * (base.py:256-306). Integer sums are exact (nchans < 2^24); float32 data are summed in
 * float64 in channel order.
 */
int spp_collapse(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps, float* out,
                 int64_t out_start, void* stream);
/*
 * spp_mask_bytes: in place, byte j of every row_bytes-long row becomes
 * (b & keep[j]) | set[j] (keep/set: device uint8 [row_bytes]). With keep/set built from a
 * channel mask and value this is kernels.mask_channels (core/kernels.py:321-345) followed by
 * the reference's pack, applied to the file's own bytes. Called by
 * Filterbank.apply_channel_mask (base.py:530-583).
 */
int spp_mask_bytes(uint8_t* data, int64_t nbytes, int64_t row_bytes, const uint8_t* keep,
                   const uint8_t* set, void* stream);
/*
 * spp_remove_zerodm: kernels.remove_zerodm (core/kernels.py:469-500) as called by
 * Filterbank.remove_zerodm (base.py:961-1028). For each sample t, z = sum_c x[t][c] and
 * out[t][c] = fma(-z, chanwts[c], x[t][c]) + bpass[c] in float64, stored as float32
 * (nbits == 32) or as the low 8 bits of its truncation (one byte per sample, unpacked).
 * chanwts, bpass: device float32 [nchans].
 */
int spp_remove_zerodm(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
                      const float* chanwts, const float* bpass, void* out, void* stream);

/*
 * Fold one gulp (kernels.fold, core/kernels.py:404-445, as called by Filterbank.fold,
 * base.py:1109-1207). The host lists the gulp's valid samples by cell key in time order:
 * key k covers samples[key_start[k] .. key_start[k+1]) and the cells key_pos[k] + b*nbins,
 * b < nbands; band b holds channels [band_start[b], band_start[b+1]). Each cell continues its
 * float32 sum (fold) and count over its samples in time order and the band's channels in
 * channel order, reading x[s + delays[c]][c] of the packed gulp (nbits 1/2/4/8 with bit order
 * `big`, or 32 for float32); a negative row wraps by nsamps. Host arrays: none; all pointers
 * are device pointers. fold/count: float32/int32 [nints * nbands * nbins]. Integer data need a
 * device workspace of spp_fold_workspace_size bytes (the gulp's band sums; none for float32).
 */
int spp_fold_workspace_size(int nbits, int64_t nchans, int64_t nsamps, int32_t nbands, int64_t* bytes);
int spp_fold(const void* in, int nbits, int big, int64_t nchans, int64_t nsamps,
             const int32_t* delays, const int32_t* samples, int64_t nlisted,
             const int32_t* key_start, const int32_t* key_pos, int32_t nkeys,
             const int32_t* band_start, int32_t nbands, int32_t nbins, float* fold, int32_t* count,
             void* workspace, int64_t workspace_bytes, void* stream);

/*
 * In-memory blocks: float32 [nrows][ncols] (channel-major FilterbankBlock data). Row r is
 * read at column (t + off[r]) mod ncols; off_host and off hold the same int64 offsets (host
 * copy for validation, device copy for the kernel), each in [0, ncols). out_cols < ncols is
 * the valid-columns form, where no offset may wrap.
 *
 * spp_roll_rows: out[r][t] = in[r][(t + off[r]) mod ncols], t < out_cols. Replaces
 * kernels.roll_block / roll_block_valid (core/kernels.py:1059-1152) as called by
 * FilterbankBlock.dedisperse (block.py:267-307) with shifts -delays.
 * spp_dmt: out[i][t] = sum over c, in channel order, float32, of in[c][(t + off[i][c]) mod
 * ncols]. Replaces kernels.dmt_block / dmt_block_valid (core/kernels.py:1155-1193) as called by
 * FilterbankBlock.dmt_transform (block.py:309-343) with shifts +delays.
 */
int spp_roll_rows(const float* in, int64_t nrows, int64_t ncols, const int64_t* off_host,
                  const int64_t* off, int64_t out_cols, float* out, void* stream);
int spp_dmt(const float* in, int64_t nchans, int64_t ncols, const int64_t* off_host,
            const int64_t* off, int64_t ndm, int64_t out_cols, float* out, void* stream);

/*
 * Running normalisation of DM series (no reference counterpart: the running mean/std
 * pre-step before MatchedFilter). For each row i of length row_len_host[i]:
 *   z[t] = (x[t] - m[t]) / s[t], m/s = mean/std over x padded symmetrically as
 *   stats.running_filter pads (core/stats.py:231-236), `window` samples, float64;
 *   s == 0 is replaced by 1 (as stats.estimate_zscore does, core/stats.py:469-471).
 * z must not alias x. workspace: spp_running_normalise_workspace_size bytes.
 * Synchronises the stream before returning.
 */
int spp_running_normalise_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                         int32_t window, int64_t* bytes);
int spp_running_normalise(const float* x, int64_t ldx, int64_t nrows,
                          const int64_t* row_len_host, int32_t window, float* z, int64_t ldz,
                          void* workspace, int64_t workspace_bytes, void* stream);

/*
 * Boxcar matched filter with argmax per row. Replaces MatchedFilter._compute
 * (core/filters.py:325-338) with temp_kind "boxcar" on zscores that are already
 * normalised (loc/scale "norm"), i.e. kernels.convolve_templates
 * (core/kernels.py:861-906) in closed form:
 *   conv_w[t] = (sum_{j<w} zt[(t+j) mod Ng] - w*S/Ng) / sqrt(w*(1 - w/Ng)),
 *   zt = circular pad of z to Ng = 5-smooth good size >= n (kernels.py:998-1005),
 *   S = sum(zt); then the first maximum over [width][t] in template-major order
 *   (np.argmax over convs, filters.py:333-338).
 * That closed form is the reference's result when Ng is even. When Ng is odd the
 * reference's irfft has its default length Ng - 1 (kernels.py:904): the row's conv is
 *   y = IDFT_{Ng-1}(Y), Y[k] = DFT_Ng(conv)[k] for 0 <= k <= (Ng-1)/2 (imaginary parts of
 *   the first and last of those bins dropped), Y[Ng-1-k] = conj(Y[k]),
 * computed in float64 through FFTs; a row of exactly Ng samples (Ng odd) cannot be filled
 * by Ng - 1 samples, and is refused with SPP_EINVAL as the reference raises.
 * widths_host: int32 [nwidths] (1 <= w <= 1024, nwidths <= 32). Outputs per row (device):
 * snr (float32 conv value at the maximum), peak (t), widx (index into widths).
 * workspace: spp_boxcar_mf_workspace_size bytes for the same row lengths and widths.
 * Synchronises the stream.
 */
int spp_boxcar_mf_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                 const int32_t* widths_host, int32_t nwidths, int64_t* bytes);
int spp_boxcar_mf(const float* z, int64_t ldz, int64_t nrows, const int64_t* row_len_host,
                  const int32_t* widths_host, int32_t nwidths, float* snr, int32_t* peak,
                  int32_t* widx, void* workspace, int64_t workspace_bytes, void* stream);
/* spp_boxcar_mf without a workspace, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops). No counterpart in the reference. Its scratch
 * is a static device array of 8 MiB per GPU held until the process ends (enough for 1024 rows
 * of 2^19 samples with the 9 widths 1..256); rows that need more are refused with
 * SPP_EINVAL. Calls no device allocator. Calls that overlap on one GPU share the scratch and
 * race: serialise them. Nothing in the scratch needs clearing (each call writes every slot
 * it reads). Synchronises the stream. */
int spp_boxcar_mf_noalloc(const float* z, int64_t ldz, int64_t nrows, const int64_t* row_len_host,
                          const int32_t* widths_host, int32_t nwidths, float* snr, int32_t* peak,
                          int32_t* widx, void* stream);

/*
 * Running normalisation and boxcar MF in one pass over the rows of x: the MatchedFilter of
 * core/filters.py:325-338 (convolve_templates, core/kernels.py:861-906, with boxcar
 * templates) applied to the running mean/std normalisation of each row
 * (core/stats.py:231-236 padding), as spp_boxcar_mf(spp_running_normalise(x)) but without
 * the normalised plane in device memory. z is computed with spp_running_normalise's
 * arithmetic; S and the box sums in float64. Each width's first maximum is taken on the
 * float64 box sums, then conv is rounded to float32 and the template-major first maximum
 * kept, so (snr, peak, widx) equal spp_boxcar_mf's except where float32 rounding makes two
 * box sums of one width equal. Rows whose good size is odd take spp_boxcar_mf's FFT path
 * (one row at a time); a row of exactly Ng (odd) samples is refused with SPP_EINVAL.
 * Row lengths 1 <= n <= min(ldx, 2^31 - 1) and n >= the largest width; window >= 1.
 * workspace: spp_normalise_boxcar_mf_workspace_size bytes. Synchronises the stream.
 */
int spp_normalise_boxcar_mf_workspace_size(int64_t nrows, const int64_t* row_len_host,
                                           const int32_t* widths_host, int32_t nwidths, int64_t* bytes);
int spp_normalise_boxcar_mf(const float* x, int64_t ldx, int64_t nrows, const int64_t* row_len_host,
                            int32_t window, const int32_t* widths_host, int32_t nwidths, float* snr,
                            int32_t* peak, int32_t* widx, void* workspace, int64_t workspace_bytes,
                            void* stream);

/*
 * Template bank convolution. Replaces kernels.convolve_templates (core/kernels.py:861-906)
 * for any template: convs[k][t] = (sum_m h_k[m] zt[(t - ref_k + m) mod Ng] - mu_k*S) /
 * sigma_k with mu_k = sum(h_k)/Ng and sigma_k = sqrt(sum(h_k^2) - sum(h_k)^2/Ng) (the
 * normalize_template of core/kernels.py:1008-1027 applied to the padded template;
 * sigma_k == 0 leaves the numerator unscaled). Computed in float64, stored float32.
 * For an odd Ng the reference's default-length irfft (kernels.py:904) is reproduced as for
 * spp_boxcar_mf, and n == Ng (odd) is refused with SPP_EINVAL.
 * temps_host: concatenated float32 template samples; temp_len_host / ref_bin_host: int32
 * [ntemps]. convs: float32 [ntemps][n]. workspace: spp_convolve_templates_workspace_size.
 */
int spp_convolve_templates_workspace_size(int64_t n, int32_t ntemps, int64_t total_taps,
                                          int64_t* bytes);
int spp_convolve_templates(const float* data, int64_t n, const float* temps_host,
                           const int32_t* temp_len_host, const int32_t* ref_bin_host,
                           int32_t ntemps, float* convs, void* workspace,
                           int64_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SPP_HIP_H */
