// One-DM dedispersion of packed 1/2/4/8-bit gulps with word-parallel (SWAR) adds: the
// dedisperse form of spp_dedisperse, used when each packed row is whole dwords.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spp {

// True when the SWAR kernel takes this launch: 1/2/4/8-bit rows of whole dwords on a 4-byte
// aligned stream.
bool gulp_swar_applies(const void* in, int nbits, int64_t nchans);

// out[out_start + t] (t < nout) = sum over channels of x[t + delays[c]][c], with x[row] for
// row < 0 taken from row + nsamps and rows >= nsamps read as 0 (only outputs t >= nout reach
// them). Blocks own `chans_per_group` channels in `nchunk` chunks (the grouping and line
// interleave of the byte-staged kernel); with groups > 1 the outputs are float atomics onto a
// zeroed output. Returns an SPP_* status.
int launch_gulp_swar(const uint8_t* in, int nbits, bool big, const int32_t* delays, float* out, int64_t nchans,
                     int64_t nsamps, int64_t nout, int64_t out_start, int64_t row_bytes, int64_t chans_per_group,
                     int64_t groups, int64_t ilv, int64_t nchunk, int use_atomic, bool compact, hipStream_t st);

// Outputs per block for nbits (the time tile).
int64_t gulp_swar_tile(int nbits);

// True when the compact form (a smaller series buffer: more blocks per CU, fewer rows per pass)
// suits a launch whose largest delay is max_delay (4/8-bit data only).
bool gulp_swar_compact(int nbits, int64_t nchans, int64_t max_delay);

// Channels per stage for nbits (a stage is a fixed run of dwords of each row).
int64_t gulp_swar_stage_channels(int nbits);

}  // namespace spp
