// Internal declarations shared by the HIP and host translation units of
// libpct.so.  Not part of the C ABI.
#pragma once

#include <cstddef>
#include <cstdint>

#include "pct/pct.h"

namespace pct {

// Largest half window the tiled kernels are instantiated for (NH <= 15).
constexpr int kMaxTiledHalfWindow = 7;

// Static description of one tiled-kernel configuration.
struct TileShape {
    int seg_width;  // lanes per y segment
    int bx, bz;     // per-lane output block
    int waves;      // wavefronts per workgroup
};

// Geometry of one tiled launch, derived from the problem and a TileShape.
struct TileGeometry {
    int hw = 0, r = 0;
    int tiles_y = 0, tiles_x = 0, tiles_z = 0;
    int group = 1, n_groups = 1;
    long long n_tiles = 0, n_items = 0;
    int lds_bytes = 0;
    int wg_per_cu = 0;
};

// Problem validation shared by every entry point.
int validate_problem(const pct_problem* pb);

// Host-side helpers implemented in pct_host.cpp.
uint64_t splitmix64(uint64_t& state);

// memcpy split over the OpenMP threads in page-aligned parts (a single
// memcpy below 4 MiB).
void parallel_copy(void* dst, const void* src, size_t bytes);

}  // namespace pct
