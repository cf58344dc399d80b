// Internal declarations shared by the HIP and host translation units of
// libpct.so.  Not part of the C ABI.
#pragma once

#include <climits>
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
    int form = 0;   // 0: squared differences; 1: cross terms where the range allows
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

// The weight divisor of a problem: pb->div when given, else
// pct_weight_divisor(sigma, beta).
double problem_divisor(const pct_problem* pb);

// Largest element count a reader accepts for all of a file's volumes together.
constexpr long long kMaxFileElements = 1LL << 34;

// Element counts of `num` volumes of rows x cols x depth read from a file,
// with every product checked for overflow before it is used: false when any
// dimension is below 1, one volume holds INT32_MAX elements or more, or the
// total exceeds kMaxFileElements.
inline bool checked_volume_elements(long long num, long long rows, long long cols,
                                    long long depth, long long* per_volume,
                                    long long* total) {
    if (num < 1 || rows < 1 || cols < 1 || depth < 1) return false;
    long long plane = 0, vol = 0, all = 0;
    if (__builtin_mul_overflow(rows, cols, &plane) || __builtin_mul_overflow(plane, depth, &vol) ||
        vol >= (long long)INT32_MAX || __builtin_mul_overflow(vol, num, &all) ||
        all > kMaxFileElements)
        return false;
    *per_volume = vol;
    *total = all;
    return true;
}

// Host-side helpers implemented in pct_host.cpp.
uint64_t splitmix64(uint64_t& state);

// memcpy split over the OpenMP threads in page-aligned parts (a single
// memcpy below 4 MiB).
void parallel_copy(void* dst, const void* src, size_t bytes);

}  // namespace pct
