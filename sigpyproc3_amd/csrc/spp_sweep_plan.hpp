// Host-side state of a sweep plan (spp_sweep_plan_create), shared by the one-plan entry points
// (sweep.hip) and the batch (batch.hip), which copies the tables of several plans.
#pragma once

#include <cstdint>
#include <vector>

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
    bool uploaded_with_g = false;  // the workspace also holds the stage-1 copy (off_g .. ws_bytes)
};
