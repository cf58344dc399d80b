// Plain FP32 multi-atlas patch fusion for gfx950 (MI355X).
//
// Semantics follow the reference fusion loop (src/CPU/NewPatchedBrainPureC.cpp:
// 1184-1212 PatchSimilarity, 1350-1449 the OpenMP loop nest; the CUDA kernels
// src/GPU/PatchBasedPseudoCT.cu:162-641 compute the same function).  The
// design is new:
//
//  * Tiled kernel (EXEC_CHOICE SM).  For a fixed atlas a and offset d the
//    27-term SSD of every target voxel is a 3x3x3 box sum of the squared
//    difference D_d(u) = (T(u) - A_a(u+d))^2, so it is computed separably:
//    x by register pairs, z by register pairs, y across lanes with two DPP
//    adds (wave_shr:1 / wave_shl:1).  Lanes map to consecutive y (coalesced
//    global and conflict-free LDS accesses); each lane owns a BZ x BX block
//    of outputs in (z, x) and keeps the scaled target patch of that block in
//    VGPRs.  Per (atlas, dz) the workgroup stages the atlas MRI planes it
//    needs (BZ+2 planes) and the CT planes (BZ planes) for its tile in LDS;
//    the (dx, dy) loop then reads LDS at compile-time offsets.  MRI values are
//    pre-scaled by sqrt(log2(e)/div) so the weight is exp2(-SSD') with no
//    multiply.  A work item is (tile, group of atlases); it writes per-group
//    float partial sums P, S, and a finalize kernel adds the groups in a fixed
//    order and applies the reference's NaN rule (cpp:1437-1443), so results
//    are deterministic.
//  * Direct kernel (EXEC_CHOICE GM / GM2).  One lane per voxel, the 27-term
//    SSD per candidate straight from global memory, FP32.  An independent
//    GPU implementation used as the GM/GM2 surface and as a cross-check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "pct_internal.h"

#include "pct_tiled.h"

namespace pct {

// Sum of the per-group partials in group order; border voxels get 0
// (the reference zeroes its output, cu:1693 / cpp:1606).
__global__ void fuse_finalize(const float* __restrict__ ws, int n_groups, long long N, int nx,
                              int ny, int nz, int r, float* __restrict__ out) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    const int z = blockIdx.z;
    if (y >= ny) return;
    const long long idx = ((long long)z * nx + x) * ny + y;
    const bool interior = z >= r && z < nz - r && x >= r && x < nx - r && y >= r && y < ny - r;
    float v = 0.f;
    if (interior) {
        float P = 0.f, S = 0.f;
        for (int g = 0; g < n_groups; ++g) {
            P += ws[(long long)(2 * g) * N + idx];
            S += ws[(long long)(2 * g + 1) * N + idx];
        }
        v = fused_value(P, S);
    }
    out[idx] = v;
}

// ---------------------------------------------------------------------------
// direct kernel (GM / GM2 surface)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void fuse_direct(const float* __restrict__ img,
                                                   const float* __restrict__ atlas,
                                                   const float* __restrict__ label, long long N,
                                                   int nx, int ny, int nz, int hw, int n_atlas,
                                                   float k, float* __restrict__ out) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    const int z = blockIdx.z;
    if (y >= ny) return;
    const int r = hw + 1;
    const long long plane = (long long)nx * ny;
    const long long idx = z * plane + (long long)x * ny + y;
    if (!(z >= r && z < nz - r && x >= r && x < nx - r && y >= r && y < ny - r)) {
        out[idx] = 0.f;
        return;
    }
    float t[27];
#pragma unroll
    for (int p = 0; p < 27; ++p)
        t[p] = img[idx + (p / 9 - 1) * plane + (long long)((p / 3) % 3 - 1) * ny + (p % 3 - 1)];
    float P = 0.f, S = 0.f;
    for (int a = 0; a < n_atlas; ++a) {
        const float* __restrict__ A = atlas + (long long)a * N;
        const float* __restrict__ L = label + (long long)a * N;
        for (int dz = -hw; dz <= hw; ++dz)
            for (int dx = -hw; dx <= hw; ++dx)
                for (int dy = -hw; dy <= hw; ++dy) {
                    const long long c = idx + dz * plane + (long long)dx * ny + dy;
                    float ssd = 0.f;
#pragma unroll
                    for (int p = 0; p < 27; ++p) {
                        const float df =
                            t[p] - A[c + (p / 9 - 1) * plane + (long long)((p / 3) % 3 - 1) * ny +
                                     (p % 3 - 1)];
                        ssd = fmaf(df, df, ssd);
                    }
                    const float w = __builtin_amdgcn_exp2f(-ssd * k);
                    P = fmaf(w, L[c], P);
                    S += w;
                }
    }
    out[idx] = fused_value(P, S);
}

// The cross-term pairs of every voxel of atlases [0, n_atlas): with AA the
// 3x3x3 sum of (scale * A)^2 (zero outside the volume) and E = 2^-AA, the
// pair (L * E, E), or (+inf, E) where the CT value L lies beyond
// kXtLabelBound or is not a number, which the slab staging flags.  A
// workgroup owns 64 y x 4 x columns of one chunk of zc planes and sweeps it
// in z: the in-plane 3x3 sums of each plane's (4+2) x (64+2) tile, staged
// through LDS, enter a running sum over three planes.  Four planes are
// loaded per step (eight loads in flight per lane, issued with the step's
// CT loads before the previous step's sums), and chunking z
// (two halo planes per chunk) gives the grid enough workgroups to hide HBM
// latency on a single subject.  Every AA is ((s(z-1) + s(z)) + s(z+1)) with
// each in-plane sum in (dx, dy) order, whatever the chunking.
constexpr int kSqBoxY = 64, kSqBoxX = 4, kSqStep = 4;
constexpr int kSqThreads = kSqBoxY * kSqBoxX;
constexpr int kSqTile = (kSqBoxX + 2) * (kSqBoxY + 2);
constexpr int kSqLoads = (kSqTile + kSqThreads - 1) / kSqThreads;
constexpr int kSqMinChunk = 8;
__global__ __launch_bounds__(kSqThreads) void atlas_sq_box(const float* __restrict__ atlas,
                                                           const float* __restrict__ label,
                                                           long long N, int nx, int ny, int nz,
                                                           int n_atlas, int zc, int n_chunks,
                                                           float scale,
                                                           f32x2* __restrict__ le) {
    __shared__ float tile[kSqStep * kSqTile];
    const int ly = threadIdx.x % kSqBoxY, lx = threadIdx.x / kSqBoxY;
    const int y0 = blockIdx.x * kSqBoxY, x0 = blockIdx.y * kSqBoxX;
    const long long plane = (long long)nx * ny;
    const int y = y0 + ly, x = x0 + lx;
    // the lane's tile elements: in-plane offset, and whether they lie in the plane
    long long off[kSqLoads];
    bool in[kSqLoads];
#pragma unroll
    for (int j = 0; j < kSqLoads; ++j) {
        const int e = threadIdx.x + j * kSqThreads;
        const int rx = e / (kSqBoxY + 2), cy = e - rx * (kSqBoxY + 2);
        const int xx = x0 - 1 + rx, yy = y0 - 1 + cy;
        in[j] = e < kSqTile && xx >= 0 && xx < nx && yy >= 0 && yy < ny;
        off[j] = in[j] ? (long long)xx * ny + yy : 0;
    }
    const int n_work = n_atlas * n_chunks;
    for (int wi = blockIdx.z; wi < n_work; wi += gridDim.z) {
        const int at = wi / n_chunks, ck = wi - at * n_chunks;
        const int zs = ck * zc, ze = min(nz, zs + zc);   // output planes [zs, ze)
        const float* __restrict__ A = atlas + (long long)at * N;
        const float* __restrict__ L = label + (long long)at * N;
        f32x2* __restrict__ out = le + (long long)at * N;
        const bool col_in = x < nx && y < ny;
        const long long col = col_in ? (long long)x * ny + y : 0;
        float s_prev = 0.f, s_cur = 0.f;   // in-plane sums of planes q-2 and q-1
        // raw values of planes p .. p+3 at the lane's tile elements, loaded
        // from clamped planes (always inside the volume) so that all of a
        // step's loads are in flight together; the step that stores them
        // zeroes the positions outside the volume and past the chunk's last
        // plane ze, and scales the rest
        float v[kSqStep][kSqLoads];
        auto load_step = [&](int p) {
#pragma unroll
            for (int k = 0; k < kSqStep; ++k) {
                const long long qo = (long long)clampi(p + k, nz) * plane;
#pragma unroll
                for (int j = 0; j < kSqLoads; ++j) v[k][j] = A[qo + off[j]];
            }
        };
        // CT values of output planes p-1 .. p+2 (plane q completes output
        // q-1), read only for the outputs the step writes
        float lab[kSqStep];
        auto load_labels = [&](int p) {
#pragma unroll
            for (int k = 0; k < kSqStep; ++k)
                lab[k] = L[(long long)clampi(p + k - 1, nz) * plane + col];
        };
        load_step(zs - 1);
        load_labels(zs - 1);
        // planes zs-1 .. ze: plane q completes output plane q-1
        for (int p = zs - 1; p <= ze; p += kSqStep) {
            __syncthreads();   // the previous step's reads of the tiles are done
            // the loads are used only here, after the barrier: the compiler
            // would otherwise move each under its own branch and wait for it
#pragma unroll
            for (int k = 0; k < kSqStep; ++k)
#pragma unroll
                for (int j = 0; j < kSqLoads; ++j) asm volatile("" : "+v"(v[k][j]));
#pragma unroll
            for (int k = 0; k < kSqStep; ++k) {
                const int q = p + k;
                const bool qin = q >= 0 && q < nz && q <= ze;
#pragma unroll
                for (int j = 0; j < kSqLoads; ++j) {
                    const int e = threadIdx.x + j * kSqThreads;
                    if (e < kSqTile) tile[k * kSqTile + e] = (qin && in[j]) ? v[k][j] * scale : 0.f;
                }
            }
            __syncthreads();
            float lab_now[kSqStep];
#pragma unroll
            for (int k = 0; k < kSqStep; ++k) lab_now[k] = lab[k];
            // the next step's loads are in flight during this step's sums
            if (p + kSqStep <= ze) {
                load_step(p + kSqStep);
                load_labels(p + kSqStep);
            }
#pragma unroll
            for (int k = 0; k < kSqStep; ++k) {
                const int q = p + k;
                if (q > ze) break;
                const float* t = tile + k * kSqTile;
                float s = 0.f;
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy) {
                        const float w = t[(lx + dx) * (kSqBoxY + 2) + ly + dy];
                        s = fmaf(w, w, s);
                    }
                if (q - 1 >= zs && col_in) {
                    const float e = __builtin_amdgcn_exp2f(-((s_prev + s_cur) + s));
                    const float l = lab_now[k];
                    out[(q - 1) * plane + col] =
                        f32x2{fabsf(l) <= kXtLabelBound ? l * e : __builtin_inff(), e};
                }
                s_prev = s_cur;
                s_cur = s;
            }
        }
    }
}

// cpp:1583-1593, float32, same operation order.
__global__ void normalize_labels_kernel(float* __restrict__ v, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        v[i] = 255.0f * ((v[i] + 1024.0f) / 4095.0f);
}

// ---------------------------------------------------------------------------
// host side: configuration table, geometry, launches
// ---------------------------------------------------------------------------

using TiledFn = void (*)(TiledArgs);

struct TiledEntry {
    TileShape shape;
    int hw;
    TiledFn fn;
    int lds_bytes;
    // resident workgroups per CU reported by the runtime (registers and LDS);
    // 0 until a gfx950 device has been seen
    mutable int occupancy;
};

template <int W, int BX, int BZ, int NW, int HW, int FORM>
TiledEntry make_entry() {
    using C = TiledCfg<W, BX, BZ, NW, HW>;
    if (FORM == 1)
        return TiledEntry{{W, BX, BZ, NW, 1}, HW, &fuse_tiled_xt<W, BX, BZ, NW, HW>,
                          XtCfg<C>::LDS_BYTES, 0};
    return TiledEntry{{W, BX, BZ, NW, 0}, HW, &fuse_tiled<W, BX, BZ, NW, HW>, C::LDS_BYTES, 0};
}

template <int W, int BX, int BZ, int NW, int FORM = 0>
void add_shape(std::vector<TiledEntry>& v) {
    v.push_back(make_entry<W, BX, BZ, NW, 1, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 2, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 3, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 4, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 5, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 6, FORM>());
    v.push_back(make_entry<W, BX, BZ, NW, 7, FORM>());
}

// The shapes compiled into the library (index = pct_tile_config.id).
const std::vector<TileShape>& shapes() {
    static const std::vector<TileShape> s = {
        {64, 4, 8, 4},   // 0: one 62-wide y segment per wave, 16 x rows per workgroup
        {32, 4, 8, 4},   // 1: two 30-wide y segments per wave, 32 x rows
        {32, 4, 4, 2},   // 2: small LDS footprint, several workgroups per CU
        {32, 2, 8, 4},   // 3: 16 x rows, 8 planes
        {64, 4, 4, 4},   // 4: 16 x rows, 4 planes
        {32, 4, 4, 4},   // 5: 32 x rows, 4 planes; two workgroups per CU up to NH=11
        {32, 2, 6, 4},   // 6: 16 x rows, 6 planes; two workgroups per CU up to NH=11
        {32, 2, 4, 4},   // 7: 16 x rows, 4 planes; three workgroups per CU up to NH=11
        {32, 4, 2, 4},   // 8: 32 x rows, 2 planes; three workgroups per CU up to NH=11
        {64, 2, 4, 4},   // 9: 8 x rows of 62-wide segments, 4 planes
        {32, 2, 2, 8},   // 10: 32 x rows, 2 planes, eight waves share one slab
        {32, 2, 2, 4},   // 11: 16 x rows, 2 planes; five workgroups per CU up to NH=11
        {32, 2, 2, 2},   // 12: 8 x rows, 2 planes, two waves; eight workgroups per CU
        {32, 2, 4, 8},   // 13: 32 x rows, 4 planes, eight waves
        {32, 4, 2, 8},   // 14: 64 x rows, 2 planes, eight waves
        // cross-term form (slab: BZ + 2 MRI planes and BZ planes of (L~, E) pairs)
        {32, 2, 4, 4, 1},   // 15: 16 x rows, 4 planes; two workgroups per CU up to NH=11
        {32, 2, 2, 4, 1},   // 16: 16 x rows, 2 planes; four workgroups per CU up to NH=11
        {32, 2, 2, 8, 1},   // 17: 32 x rows, 2 planes, eight waves
        {32, 4, 2, 4, 1},   // 18: 32 x rows, 2 planes
        {64, 2, 2, 4, 1},   // 19: 8 x rows of 62-wide segments, 2 planes
    };
    return s;
}

static const std::vector<TiledEntry>& entries() {
    static const std::vector<TiledEntry> e = [] {
        std::vector<TiledEntry> v;
        add_shape<64, 4, 8, 4>(v);
        add_shape<32, 4, 8, 4>(v);
        add_shape<32, 4, 4, 2>(v);
        add_shape<32, 2, 8, 4>(v);
        add_shape<64, 4, 4, 4>(v);
        add_shape<32, 4, 4, 4>(v);
        add_shape<32, 2, 6, 4>(v);
        add_shape<32, 2, 4, 4>(v);
        add_shape<32, 4, 2, 4>(v);
        add_shape<64, 2, 4, 4>(v);
        add_shape<32, 2, 2, 8>(v);
        add_shape<32, 2, 2, 4>(v);
        add_shape<32, 2, 2, 2>(v);
        add_shape<32, 2, 4, 8>(v);
        add_shape<32, 4, 2, 8>(v);
        add_shape<32, 2, 4, 4, 1>(v);
        add_shape<32, 2, 2, 4, 1>(v);
        add_shape<32, 2, 2, 8, 1>(v);
        add_shape<32, 4, 2, 4, 1>(v);
        add_shape<64, 2, 2, 4, 1>(v);
        return v;
    }();
    return e;
}

static const TiledEntry* find_entry(int shape_id, int hw) {
    if (shape_id < 0 || shape_id >= (int)shapes().size()) return nullptr;
    if (hw < 1 || hw > kMaxTiledHalfWindow) return nullptr;
    const auto& e = entries()[shape_id * kMaxTiledHalfWindow + (hw - 1)];
    return &e;
}

int validate_problem(const pct_problem* pb) {
    if (!pb) return PCT_ERR_INVALID;
    if (pb->nx < 1 || pb->ny < 1 || pb->nz < 1 || pb->n_atlas < 1) return PCT_ERR_INVALID;
    if (pb->nh < 1 || pb->patch != 3) return PCT_ERR_INVALID;
    if (!(pb->div >= 0.0) || !std::isfinite(pb->div)) return PCT_ERR_INVALID;
    if (pb->div == 0.0) {
        if (!(pb->sigma > 0.0) || !(pb->beta > 0.0)) return PCT_ERR_INVALID;
        if (!std::isfinite(pb->sigma) || !std::isfinite(pb->beta)) return PCT_ERR_INVALID;
    }
    const long long N = (long long)pb->nx * pb->ny * pb->nz;
    if (pb->nx > 65535 || pb->ny > 65535 || pb->nz > 65535) return PCT_ERR_TOO_LARGE;
    if (N * (long long)pb->n_atlas > (1LL << 40)) return PCT_ERR_TOO_LARGE;
    return PCT_OK;
}

double problem_divisor(const pct_problem* pb) {
    return pb->div > 0.0 ? pb->div : pct_weight_divisor(pb->sigma, pb->beta);
}

static int interior(int n, int r) { return std::max(0, n - 2 * r); }

bool make_geometry(const pct_problem* pb, int shape_id, int group, TileGeometry* g) {
    const int hw = (pb->nh - 1) / 2;
    const TiledEntry* e = find_entry(shape_id, hw);
    if (!e) return false;
    const TileShape& s = e->shape;
    if (e->lds_bytes > kLdsPerCu) return false;
    const int S = 64 / s.seg_width;
    const int WX = s.waves * S * s.bx;
    const int OY = s.seg_width - 2;
    g->hw = hw;
    g->r = hw + 1;
    const int iy = interior(pb->ny, g->r), ix = interior(pb->nx, g->r), iz = interior(pb->nz, g->r);
    g->tiles_y = (iy + OY - 1) / OY;
    g->tiles_x = (ix + WX - 1) / WX;
    g->tiles_z = (iz + s.bz - 1) / s.bz;
    g->n_tiles = (long long)g->tiles_y * g->tiles_x * g->tiles_z;
    g->group = std::max(1, std::min(group, pb->n_atlas));
    g->n_groups = (pb->n_atlas + g->group - 1) / g->group;
    g->n_items = g->n_tiles * g->n_groups;
    g->lds_bytes = e->lds_bytes;
    const int by_lds = kLdsPerCu / e->lds_bytes;
    const int by_waves = 32 / s.waves;  // 8 waves per SIMD slot budget, 4 SIMDs
    g->wg_per_cu = e->occupancy > 0 ? e->occupancy : std::max(1, std::min(by_lds, by_waves));
    return g->n_items < INT_MAX;
}

// Cost model: issue slots per output-candidate for one lane, including the
// halo lanes and halo rows, times the number of rounds the grid needs.
// Relative VALU issue rate a SIMD sustains on this kernel with w resident
// waves.  Fitted to the tiled kernels measured on MI355X at 222x222x112
// (n=10, NH=11; n=5, NH=9): one wave per SIMD issues far below the VALU
// rate, throughput grows with occupancy and saturates around five waves.
static double simd_efficiency(double w) {
    if (w <= 1.0) return 0.25 * w;
    if (w <= 2.0) return 0.25 + 0.30 * (w - 1.0);
    if (w <= 3.0) return 0.55 + 0.13 * (w - 2.0);
    if (w <= 4.0) return 0.68 + 0.04 * (w - 3.0);
    return std::min(0.78, 0.72 + 0.06 * (w - 4.0));
}

// Cost model: VALU issue slots of one item per wave, times the rounds of
// items the grid needs, divided by the efficiency the occupancy allows.
// Per (dx, dy) step the instruction count is the fit
//   1.7 (BZ+2)(BX+2) + 1.5 (BZ+2) BX + 6.5 BZ BX
// (squared differences, x sums, then z sums, DPP y sums, exp, accumulation
// per output), within 5% of the compiled kernels; exp issues as four slots.
double model_item_time(const pct_problem* pb, int shape_id, int group, TileGeometry* geo) {
    TileGeometry g;
    if (!make_geometry(pb, shape_id, group, &g)) return -1;
    const TileShape& s = shapes()[shape_id];
    const double bx = s.bx, bz = s.bz;
    const double nh = 2 * g.hw + 1;
    // cross-term shapes: per step 2 (BZ+2) BX FMA-chain instructions for the
    // x sums, 1.5 BZ BX for z, 2 BZ BX DPP adds for y, 2 BZ BX FMAs for P and
    // S, BZ BX exp and one address add; scaled like the fit below (measured:
    // {32, 2, 2, 4} at 222x222x112, n=10, NH=11 runs at 85 % of its VALU
    // issue rate, its squared-difference twin at 100 %)
    const double step = s.form == 1
                            ? 1.28 / 0.85 * (2.0 * (bz + 2) * bx + 6.5 * bz * bx + 1.0)
                            : 1.7 * (bz + 2) * (bx + 2) + 1.5 * (bz + 2) * bx + 6.5 * bz * bx +
                                  3.0 * bz * bx;
    // staging per dz step: address, load, scale, LDS store per element (the
    // cross-term slab also loads AA and forms (L~, E) for its CT planes)
    const int S = 64 / s.seg_width;
    const double rows = (s.form == 1 ? 3 * bz + 2 : 2 * bz + 2) *
                        (s.waves * S * s.bx + 2 + 2 * g.hw);
    const double stage = rows / s.waves * 8.0 * ((s.seg_width + 2 * g.hw + 63) / 64);
    // one-segment (64-lane) shapes measured about 1.5x slower than their
    // instruction counts predict
    const double width_penalty = s.seg_width == 64 ? 1.5 : 1.0;
    const double wave_time = width_penalty * g.group * nh * (nh * nh * step + stage);
    const double w = (double)g.wg_per_cu * s.waves / 4.0;  // waves per SIMD
    if (geo) *geo = g;
    return wave_time * std::max(1.0, w) / simd_efficiency(w);
}

static double model_cost(const pct_problem* pb, int shape_id, int group, TileGeometry* geo) {
    TileGeometry g;
    const double item_time = model_item_time(pb, shape_id, group, &g);
    if (item_time < 0) return -1;
    const double rounds = std::ceil((double)g.n_items / ((double)kNumCu * g.wg_per_cu));
    if (geo) *geo = g;
    return rounds * item_time;
}

// Tile shape that measured fastest on MI355X for a half window, or -1 where
// the cost model decides.  Autotuning at 222x222x112 (n=10, NH=11; n=5 and
// 7, NH=11), 192x192x96 and 256x256x128 (NH=9), 160x160x96 and 128x128x64
// (NH=7): the analytic model is within about 15 % of the measured times,
// while the best shapes of a window lie within 5 % of each other.
int preferred_shape(int hw) {
    switch (hw) {
        // every shape timed at group 1 and 2 on one box (ms): NH 7: the
        // cross-term {32, 2, 2, 4} (16) 0.318 against 0.312 for {64, 2, 2, 4}
        // (19) at 160x160x96, n=1, and 0.928 against 0.946 at 128x128x64,
        // n=10; NH 9: 19 2.124 against 2.366 for 16 at 192x192x96, n=3, and
        // 16.65 against 17.03 at 256x256x128, n=10; NH 11: 16 8.93 and
        // 17.13 at 222x222x112, n=5 and 10, against 9.03 and 17.79 for
        // {32, 2, 2, 8} (17) and 9.85 and 19.62 for the best
        // squared-difference shape
        case 3: return 16;
        case 4: return 19;
        case 5: return 16;
        default: return -1;
    }
}

struct ShapeKey {
    int nx, ny, nz, n, nh;
    bool operator<(const ShapeKey& o) const {
        return std::tie(nx, ny, nz, n, nh) < std::tie(o.nx, o.ny, o.nz, o.n, o.nh);
    }
};

static std::mutex g_cfg_mutex;
static std::map<ShapeKey, std::pair<int, int>> g_tuned;   // shape -> (shape_id, group)
static int g_forced_id = -1, g_forced_group = 0;

static bool choose(const pct_problem* pb, int* shape_id, int* group, double* cost,
                   TileGeometry* geo) {
    {
        std::lock_guard<std::mutex> lk(g_cfg_mutex);
        if (g_forced_id >= 0) {
            *shape_id = g_forced_id;
            *group = g_forced_group > 0 ? g_forced_group : 1;
            double c = model_cost(pb, *shape_id, *group, geo);
            if (cost) *cost = c;
            return c >= 0;
        }
        auto it = g_tuned.find(ShapeKey{pb->nx, pb->ny, pb->nz, pb->n_atlas, pb->nh});
        if (it != g_tuned.end()) {
            *shape_id = it->second.first;
            *group = it->second.second;
            double c = model_cost(pb, *shape_id, *group, geo);
            if (cost) *cost = c;
            return c >= 0;
        }
    }
    // a measured preference for the window fixes the shape and the model
    // picks the atlas group; without one (or where it does not fit) the model
    // picks both
    const int pref = preferred_shape((pb->nh - 1) / 2);
    double best = -1;
    for (int pass = pref >= 0 ? 0 : 1; pass < 2 && best < 0; ++pass)
        for (int sid = 0; sid < (int)shapes().size(); ++sid) {
            if (pass == 0 && sid != pref) continue;
            for (int grp = 1; grp <= pb->n_atlas; ++grp) {
                TileGeometry g;
                const double c = model_cost(pb, sid, grp, &g);
                if (c < 0) continue;
                if (best < 0 || c < best * 0.999) {
                    best = c;
                    *shape_id = sid;
                    *group = grp;
                    if (geo) *geo = g;
                }
            }
        }
    if (cost) *cost = best;
    return best >= 0;
}

// ---------------------------------------------------------------------------
// device / workspace management
// ---------------------------------------------------------------------------

struct DeviceState {
    bool checked = false;
    bool ok = false;
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

static std::mutex g_dev_mutex;
static std::map<int, DeviceState> g_dev;

int check_device(int* dev_out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PCT_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    DeviceState& st = g_dev[dev];
    if (!st.checked) {
        hipDeviceProp_t prop;
        st.checked = true;
        st.ok = hipGetDeviceProperties(&prop, dev) == hipSuccess &&
                std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
        if (st.ok) {
            for (const auto& e : entries()) {
                if (e.lds_bytes > kLdsPerCu) continue;
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(e.fn),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          e.lds_bytes);
                int n = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &n, reinterpret_cast<const void*>(e.fn), e.shape.waves * 64,
                        (size_t)e.lds_bytes) == hipSuccess && n > 0)
                    e.occupancy = n;
                else
                    (void)hipGetLastError();
            }
        }
    }
    if (dev_out) *dev_out = dev;
    return st.ok ? PCT_OK : PCT_ERR_NO_DEVICE;
}

int ensure_workspace(int dev, size_t bytes, void** ptr) {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    DeviceState& st = g_dev[dev];
    if (st.ws_bytes < bytes) {
        if (st.ws) {
            if (hipDeviceSynchronize() != hipSuccess) return PCT_ERR_HIP;
            (void)hipFree(st.ws);
            st.ws = nullptr;
            st.ws_bytes = 0;
        }
        if (hipMalloc(&st.ws, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return PCT_ERR_NO_MEMORY;
        }
        st.ws_bytes = bytes;
    }
    *ptr = st.ws;
    return PCT_OK;
}

bool is_xterm_shape(int shape_id) {
    return shape_id >= 0 && shape_id < (int)shapes().size() && shapes()[shape_id].form == 1;
}

size_t tiled_le_offset(const pct_problem* pb, int n_groups) {
    const size_t N = (size_t)pb->nx * pb->ny * pb->nz;
    return ((2 * (size_t)n_groups * N * sizeof(float)) + 255) & ~(size_t)255;
}

size_t tiled_ws_bytes(const pct_problem* pb, int n_groups, int shape_id) {
    const size_t N = (size_t)pb->nx * pb->ny * pb->nz;
    size_t bytes = tiled_le_offset(pb, n_groups);
    if (is_xterm_shape(shape_id))
        bytes += ((size_t)pb->n_atlas * N * sizeof(f32x2) + 255) & ~(size_t)255;
    return bytes;
}

int launch_atlas_pairs(const pct_problem* pb, const float* atlas, const float* label,
                       int n_atlas, float scale, f32x2* le, hipStream_t stream) {
    if (n_atlas <= 0) return PCT_OK;
    const long long N = (long long)pb->nx * pb->ny * pb->nz;
    const int gy = (pb->ny + kSqBoxY - 1) / kSqBoxY, gx = (pb->nx + kSqBoxX - 1) / kSqBoxX;
    // z chunks: about 8192 workgroups (eight per CU, four waves each), chunks
    // of at least kSqMinChunk planes
    const long long base = (long long)gy * gx * n_atlas;
    int n_chunks = (int)std::min<long long>((8192 + base - 1) / base,
                                            std::max(1, pb->nz / kSqMinChunk));
    n_chunks = std::max(1, n_chunks);
    const int zc = (pb->nz + n_chunks - 1) / n_chunks;
    n_chunks = (pb->nz + zc - 1) / zc;
    const dim3 grid(gy, gx, (unsigned)std::min<long long>((long long)n_atlas * n_chunks, 65535));
    hipLaunchKernelGGL(atlas_sq_box, grid, dim3(kSqThreads), 0, stream, atlas, label, N, pb->nx,
                       pb->ny, pb->nz, n_atlas, zc, n_chunks, scale, le);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

static bool valid_choice(int choice) {
    return choice == PCT_CHOICE_GM || choice == PCT_CHOICE_GM2 || choice == PCT_CHOICE_SM ||
           choice == PCT_CHOICE_DIRECT;
}

// GM, GM2 and SM run the tiled kernel where it is instantiated; DIRECT and
// other window sizes run the direct kernel.
static bool is_tiled_choice(const pct_problem* pb, int choice) {
    return choice != PCT_CHOICE_DIRECT && (pb->nh - 1) / 2 <= kMaxTiledHalfWindow &&
           (pb->nh - 1) / 2 >= 1;
}

int launch_finalize(const pct_problem* pb, const TileGeometry& g, const float* ws, float* out,
                    hipStream_t stream) {
    const long long N = (long long)pb->nx * pb->ny * pb->nz;
    const dim3 fb(256), fgrid((pb->ny + 255) / 256, pb->nx, pb->nz);
    hipLaunchKernelGGL(fuse_finalize, fgrid, fb, 0, stream, ws, g.n_items > 0 ? g.n_groups : 0, N,
                       pb->nx, pb->ny, pb->nz, g.r, out);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

TiledArgs make_tiled_args(const pct_problem* pb, const TileGeometry& g, const float* img,
                          const float* atlas, const float* label, float* ws) {
    TiledArgs a;
    a.img = img;
    a.atlas = atlas;
    a.label = label;
    a.ws = ws;
    a.N = (long long)pb->nx * pb->ny * pb->nz;
    a.nx = pb->nx;
    a.ny = pb->ny;
    a.nz = pb->nz;
    a.r = g.r;
    a.tiles_y = g.tiles_y;
    a.tiles_x = g.tiles_x;
    a.n_tiles = (int)g.n_tiles;
    a.n_items = (int)g.n_items;
    a.n_atlas = pb->n_atlas;
    a.group = g.group;
    const double div = problem_divisor(pb);
    a.scale = (float)std::sqrt(1.4426950408889634 / div);
    a.only = nullptr;
    a.le = nullptr;
    return a;
}

// Launches the single-problem tiled kernel of `shape_id` over the items
// described by `a` (no finalize pass).
int launch_tiled_items(const TiledArgs& a, int shape_id, int hw, hipStream_t stream) {
    const TiledEntry* e = find_entry(shape_id, hw);
    if (!e) return PCT_ERR_UNSUPPORTED;
    if (is_xterm_shape(shape_id) && !a.le) return PCT_ERR_INVALID;   // pair pass not run
    if (a.n_items <= 0) return PCT_OK;
    TiledArgs copy = a;
    void* kargs[] = {&copy};
    if (hipLaunchKernel(reinterpret_cast<const void*>(e->fn), dim3((unsigned)a.n_items),
                        dim3(e->shape.waves * 64), kargs, (size_t)e->lds_bytes,
                        stream) != hipSuccess)
        return PCT_ERR_HIP;
    return PCT_OK;
}

// Optional per-launch HIP events around the fusion kernels and the pruned
// kernel (see pct_kernel_timing).  Events are created once and reused.
struct KernelTimer {
    std::mutex mu;
    bool enabled = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
    static constexpr size_t kMax = 4096;

    // Returns the event pair to record around the next launch, or nulls.
    std::pair<hipEvent_t, hipEvent_t> next() {
        std::lock_guard<std::mutex> lk(mu);
        if (!enabled || used >= kMax) return {nullptr, nullptr};
        if (used == pool.size()) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
                return {nullptr, nullptr};
            pool.emplace_back(a, b);
        }
        return pool[used++];
    }
};
static KernelTimer g_timer;

TimedLaunch::TimedLaunch(hipStream_t st) : s(st) {
    const auto ev = g_timer.next();
    begin = ev.first;
    end = ev.second;
    if (begin) (void)hipEventRecord(begin, s);
}

TimedLaunch::~TimedLaunch() {
    if (end) (void)hipEventRecord(end, s);
}

static int launch_tiled(const pct_problem* pb, const float* img, const float* atlas,
                        const float* label, float* out, void* ws, size_t ws_bytes,
                        int shape_id, const TileGeometry& g, hipStream_t stream) {
    const TiledEntry* e = find_entry(shape_id, g.hw);
    if (!e) return PCT_ERR_UNSUPPORTED;
    if (ws_bytes < tiled_ws_bytes(pb, g.n_groups, shape_id)) return PCT_ERR_INVALID;
    TiledArgs a = make_tiled_args(pb, g, img, atlas, label, static_cast<float*>(ws));
    if (is_xterm_shape(shape_id) && g.n_items > 0) {
        f32x2* le = reinterpret_cast<f32x2*>(static_cast<char*>(ws) +
                                             tiled_le_offset(pb, g.n_groups));
        const int rc = launch_atlas_pairs(pb, atlas, label, pb->n_atlas, a.scale, le, stream);
        if (rc) return rc;
        a.le = le;
    }
    if (g.n_items > 0) {
        void* kargs[] = {&a};
        TimedLaunch timed(stream);
        if (hipLaunchKernel(reinterpret_cast<const void*>(e->fn), dim3((unsigned)g.n_items),
                            dim3(e->shape.waves * 64), kargs, (size_t)e->lds_bytes,
                            stream) != hipSuccess)
            return PCT_ERR_HIP;
    }
    return launch_finalize(pb, g, static_cast<const float*>(ws), out, stream);
}

int launch_direct(const pct_problem* pb, const float* img, const float* atlas,
                         const float* label, float* out, hipStream_t stream) {
    const long long N = (long long)pb->nx * pb->ny * pb->nz;
    const int hw = (pb->nh - 1) / 2;
    const double div = problem_divisor(pb);
    const float k = (float)(1.4426950408889634 / div);
    const dim3 block(256), grid((pb->ny + 255) / 256, pb->nx, pb->nz);
    TimedLaunch timed(stream);
    hipLaunchKernelGGL(fuse_direct, grid, block, 0, stream, img, atlas, label, N, pb->nx, pb->ny,
                       pb->nz, hw, pb->n_atlas, k, out);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

}  // namespace pct

extern "C" int pct_kernel_timing(int enable) {
    pct::KernelTimer& t = pct::g_timer;
    std::lock_guard<std::mutex> lk(t.mu);
    t.enabled = enable != 0;
    if (t.enabled) t.used = 0;
    return PCT_OK;
}

extern "C" int pct_kernel_timing_read(double* total_ms, int* launches) {
    if (!total_ms || !launches) return PCT_ERR_INVALID;
    pct::KernelTimer& t = pct::g_timer;
    std::lock_guard<std::mutex> lk(t.mu);
    double sum = 0;
    for (size_t i = 0; i < t.used; ++i) {
        if (hipEventSynchronize(t.pool[i].second) != hipSuccess) return PCT_ERR_HIP;
        float ms = 0;
        if (hipEventElapsedTime(&ms, t.pool[i].first, t.pool[i].second) != hipSuccess)
            return PCT_ERR_HIP;
        sum += ms;
    }
    *total_ms = sum;
    *launches = (int)t.used;
    return PCT_OK;
}

// ---------------------------------------------------------------------------
// C ABI (GPU part)
// ---------------------------------------------------------------------------

using namespace pct;

extern "C" int pct_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

extern "C" int pct_set_device(int device) {
    if (device < 0 || device >= pct_device_count()) return PCT_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return PCT_ERR_NO_DEVICE;
    }
    return check_device(nullptr);
}

extern "C" int pct_normalize_labels_dev(float* d_label, size_t n, pct_stream_t stream) {
    if (!d_label && n) return PCT_ERR_INVALID;
    int rc = check_device(nullptr);
    if (rc) return rc;
    if (n == 0) return PCT_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(normalize_labels_kernel, dim3(blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_label, (long long)n);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

extern "C" size_t pct_fuse_workspace_size(const pct_problem* pb, int choice) {
    if (validate_problem(pb) != PCT_OK) return 0;
    if (!is_tiled_choice(pb, choice)) return 0;
    int sid = 0, grp = 1;
    TileGeometry g;
    if (!choose(pb, &sid, &grp, nullptr, &g)) return 0;
    return tiled_ws_bytes(pb, g.n_groups, sid);
}

static int fuse_dev_impl(const pct_problem* pb, const float* img, const float* atlas,
                         const float* label, float* out, int choice, void* ws, size_t ws_bytes,
                         bool own_ws, hipStream_t stream) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!img || !atlas || !label || !out) return PCT_ERR_INVALID;
    if (!valid_choice(choice)) return PCT_ERR_INVALID;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    if (!is_tiled_choice(pb, choice)) return launch_direct(pb, img, atlas, label, out, stream);
    int sid = 0, grp = 1;
    TileGeometry g;
    if (!choose(pb, &sid, &grp, nullptr, &g)) return PCT_ERR_UNSUPPORTED;
    const size_t need = tiled_ws_bytes(pb, g.n_groups, sid);
    if (own_ws) {
        rc = ensure_workspace(dev, need, &ws);
        if (rc) return rc;
        ws_bytes = need;
    } else if (!ws || ws_bytes < need || (reinterpret_cast<uintptr_t>(ws) & 255)) {
        return PCT_ERR_INVALID;
    }
    return launch_tiled(pb, img, atlas, label, out, ws, ws_bytes, sid, g, stream);
}

extern "C" int pct_fuse_dev(const pct_problem* pb, const float* d_img, const float* d_atlas,
                            const float* d_label, float* d_out, int choice, pct_stream_t stream) {
    return fuse_dev_impl(pb, d_img, d_atlas, d_label, d_out, choice, nullptr, 0, true,
                         reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_fuse_dev_ws(const pct_problem* pb, const float* d_img, const float* d_atlas,
                               const float* d_label, float* d_out, int choice, void* d_ws,
                               size_t ws_bytes, pct_stream_t stream) {
    return fuse_dev_impl(pb, d_img, d_atlas, d_label, d_out, choice, d_ws, ws_bytes, false,
                         reinterpret_cast<hipStream_t>(stream));
}

// ---- allocation-free variant of pct_fuse_dev --------------------------------
//
// The partial sums live in `scratch`, a file-scope device array of fixed size
// that every GPU holds from module load to process end.  Its address on a
// device comes from hipGetSymbolAddress (the array's name in host code is
// only a stand-in), fetched once per device with that device current.  The
// run uses one atlas group (every item sums all atlases, so the partials are
// 2 N floats and every interior voxel is written by exactly one item before
// the finalize pass reads it: nothing needs clearing) and the
// squared-difference kernels, which need no cross-term pairs.

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mutex;
static std::map<int, void*> g_scratch_ptr;   // device -> address of `scratch`

static int scratch_address(int dev, void** ptr) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        // `dev` is current: check_device read it with hipGetDevice
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return PCT_ERR_HIP;
        }
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *ptr = it->second;
    return PCT_OK;
}

// The squared-difference shape with one atlas group: the window's measured
// squared-difference preference where it fits, else the cost model's choice
// among squared-difference shapes.
static bool choose_noalloc(const pct_problem* pb, int* shape_id, TileGeometry* geo) {
    static const int kSquaredPref[kMaxTiledHalfWindow + 1] = {-1, -1, -1, 11, 9, 7, -1, -1};
    const int hw = (pb->nh - 1) / 2;
    const int pref = hw >= 1 && hw <= kMaxTiledHalfWindow ? kSquaredPref[hw] : -1;
    if (pref >= 0 && make_geometry(pb, pref, pb->n_atlas, geo)) {
        *shape_id = pref;
        return true;
    }
    double best = -1;
    for (int sid = 0; sid < (int)shapes().size(); ++sid) {
        if (is_xterm_shape(sid)) continue;
        TileGeometry g;
        const double c = model_cost(pb, sid, pb->n_atlas, &g);
        if (c >= 0 && (best < 0 || c < best * 0.999)) {
            best = c;
            *shape_id = sid;
            *geo = g;
        }
    }
    return best >= 0;
}

extern "C" int pct_fuse_dev_noalloc(const pct_problem* pb, const float* d_img, const float* d_atlas,
                                    const float* d_label, float* d_out, int choice,
                                    pct_stream_t stream) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!d_img || !d_atlas || !d_label || !d_out) return PCT_ERR_INVALID;
    if (!valid_choice(choice)) return PCT_ERR_INVALID;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!is_tiled_choice(pb, choice))   // the direct kernel uses no scratch
        return launch_direct(pb, d_img, d_atlas, d_label, d_out, s);
    int sid = -1;
    TileGeometry g;
    if (!choose_noalloc(pb, &sid, &g)) return PCT_ERR_UNSUPPORTED;
    const size_t need = tiled_ws_bytes(pb, g.n_groups, sid);
    if (need > kNoAllocScratchBytes) return PCT_ERR_TOO_LARGE;
    void* ws = nullptr;
    rc = scratch_address(dev, &ws);
    if (rc) return rc;
    TiledArgs a = make_tiled_args(pb, g, d_img, d_atlas, d_label, static_cast<float*>(ws));
    rc = launch_tiled_items(a, sid, g.hw, s);
    if (rc) return rc;
    return launch_finalize(pb, g, static_cast<const float*>(ws), d_out, s);
}

// Device buffers, streams and events of the host-buffer entry points, kept
// per device between calls (grown on demand, freed by pct_release).  One
// host-buffer call at a time runs on a device.
constexpr int kCopyChunks = 10;  // atlas chunks whose copies overlap the fusion
// Host buffers are staged through a ring of pinned buffers: the host
// threads copy one piece into a pinned buffer while the DMA engine moves
// the earlier pieces, which is about twice the rate of a copy from pageable
// memory.
constexpr int kPinBufs = 4;
constexpr size_t kPinBytes = (size_t)16 << 20;
struct HostCallState {
    std::mutex mu;
    float *img = nullptr, *atlas = nullptr, *label = nullptr, *out = nullptr;
    size_t n_cap = 0, na_cap = 0;
    // workspace of the host-buffer calls (partial sums, cross-term pairs, post-chain
    // lists): their own, so that they never share memory with device-buffer
    // calls still running on the caller's streams
    void* ws = nullptr;
    size_t ws_cap = 0;
    // chunk kernels alternate between two compute streams, so that the tail
    // of one chunk overlaps the start of the next
    hipStream_t compute = nullptr, compute2 = nullptr, copy = nullptr;
    hipEvent_t ev[kCopyChunks] = {};
    hipEvent_t join = nullptr;
    void* pin[kPinBufs] = {};          // pinned staging ring (all or none)
    hipEvent_t pin_ev[kPinBufs] = {};  // last transfer that used each buffer
    bool pin_busy[kPinBufs] = {};
    int pin_next = 0;
};
static std::mutex g_host_mutex;
static std::map<int, HostCallState*> g_host_state;

static HostCallState& host_state(int dev) {
    std::lock_guard<std::mutex> lk(g_host_mutex);
    HostCallState*& st = g_host_state[dev];
    if (!st) st = new HostCallState();
    return *st;
}

static void free_pinned(HostCallState& st) {
    for (int b = 0; b < kPinBufs; ++b) {
        if (st.pin[b]) (void)hipHostFree(st.pin[b]);
        if (st.pin_ev[b]) (void)hipEventDestroy(st.pin_ev[b]);
        st.pin[b] = nullptr;
        st.pin_ev[b] = nullptr;
    }
}

// Host -> device through the pinned ring, in pieces of kPinBytes on stream
// cs.  Without the ring (pinned memory refused) a plain pageable copy.
static int stage_h2d(HostCallState& st, void* d_dst, const void* h_src, size_t bytes,
                     hipStream_t cs) {
    if (!st.pin[0])
        return hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, cs) == hipSuccess
                   ? PCT_OK
                   : PCT_ERR_HIP;
    for (size_t o = 0; o < bytes; o += kPinBytes) {
        const size_t len = std::min(kPinBytes, bytes - o);
        const int b = st.pin_next;
        st.pin_next = (b + 1) % kPinBufs;
        if (st.pin_busy[b] && hipEventSynchronize(st.pin_ev[b]) != hipSuccess) return PCT_ERR_HIP;
        parallel_copy(st.pin[b], static_cast<const char*>(h_src) + o, len);
        if (hipMemcpyAsync(static_cast<char*>(d_dst) + o, st.pin[b], len, hipMemcpyHostToDevice,
                           cs) != hipSuccess ||
            hipEventRecord(st.pin_ev[b], cs) != hipSuccess)
            return PCT_ERR_HIP;
        st.pin_busy[b] = true;
    }
    return PCT_OK;
}

// Device -> host through the pinned ring on stream s: up to kPinBufs pieces
// in flight, each copied out by the host threads as it lands.  Returns when
// the host buffer holds the data.
static int stage_d2h(HostCallState& st, void* h_dst, const void* d_src, size_t bytes,
                     hipStream_t s) {
    if (!st.pin[0]) {
        if (hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return PCT_ERR_HIP;
        return PCT_OK;
    }
    const size_t pieces = (bytes + kPinBytes - 1) / kPinBytes;
    auto piece_len = [&](size_t i) { return std::min(kPinBytes, bytes - i * kPinBytes); };
    auto drain = [&](size_t i) {
        const int b = (int)(i % kPinBufs);
        if (hipEventSynchronize(st.pin_ev[b]) != hipSuccess) return PCT_ERR_HIP;
        parallel_copy(static_cast<char*>(h_dst) + i * kPinBytes, st.pin[b], piece_len(i));
        return PCT_OK;
    };
    for (size_t i = 0; i < pieces; ++i) {
        const int b = (int)(i % kPinBufs);
        if (i >= (size_t)kPinBufs) {
            const int rc = drain(i - kPinBufs);
            if (rc) return rc;
        } else if (st.pin_busy[b] && hipEventSynchronize(st.pin_ev[b]) != hipSuccess) {
            return PCT_ERR_HIP;
        }
        if (hipMemcpyAsync(st.pin[b], static_cast<const char*>(d_src) + i * kPinBytes,
                           piece_len(i), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(st.pin_ev[b], s) != hipSuccess)
            return PCT_ERR_HIP;
        st.pin_busy[b] = true;
    }
    for (size_t i = pieces > (size_t)kPinBufs ? pieces - kPinBufs : 0; i < pieces; ++i) {
        const int rc = drain(i);
        if (rc) return rc;
    }
    return PCT_OK;
}

static void free_host_buffers(HostCallState& st) {
    if (st.img) (void)hipFree(st.img);
    if (st.atlas) (void)hipFree(st.atlas);
    if (st.label) (void)hipFree(st.label);
    if (st.out) (void)hipFree(st.out);
    if (st.ws) (void)hipFree(st.ws);
    st.img = st.atlas = st.label = st.out = nullptr;
    st.ws = nullptr;
    st.n_cap = st.na_cap = st.ws_cap = 0;
}

// Grows the host-call workspace to `bytes` (the previous call drained every
// stream, so nothing still uses it).
static int host_workspace(HostCallState& st, size_t bytes, void** ws) {
    if (st.ws_cap < bytes) {
        if (st.ws) (void)hipFree(st.ws);
        st.ws = nullptr;
        st.ws_cap = 0;
        if (hipMalloc(&st.ws, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return PCT_ERR_NO_MEMORY;
        }
        st.ws_cap = bytes;
    }
    *ws = st.ws;
    return PCT_OK;
}

static int prepare_host_state(HostCallState& st, size_t N, size_t NA) {
    if (!st.compute) {
        if (hipStreamCreateWithFlags(&st.compute, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&st.compute2, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&st.copy, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&st.join, hipEventDisableTiming) != hipSuccess)
            return PCT_ERR_HIP;
        for (auto& e : st.ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return PCT_ERR_HIP;
        bool pinned = true;
        for (int b = 0; b < kPinBufs && pinned; ++b)
            pinned = hipHostMalloc(&st.pin[b], kPinBytes, hipHostMallocDefault) == hipSuccess &&
                     hipEventCreateWithFlags(&st.pin_ev[b], hipEventDisableTiming) == hipSuccess;
        if (!pinned) {
            // the pageable copies still work, only slower
            (void)hipGetLastError();
            free_pinned(st);
        }
    }
    // the previous call drained every stream
    for (int b = 0; b < kPinBufs; ++b) st.pin_busy[b] = false;
    st.pin_next = 0;
    if (st.n_cap < N || st.na_cap < NA) {
        free_host_buffers(st);   // the workspace too; host_workspace regrows it
        if (hipMalloc(&st.img, N * 4) != hipSuccess || hipMalloc(&st.atlas, NA * 4) != hipSuccess ||
            hipMalloc(&st.label, NA * 4) != hipSuccess || hipMalloc(&st.out, N * 4) != hipSuccess) {
            (void)hipGetLastError();
            free_host_buffers(st);
            return PCT_ERR_NO_MEMORY;
        }
        st.n_cap = N;
        st.na_cap = NA;
    }
    return PCT_OK;
}

// Host buffers in and out around the device fusion (cu:1687-1744); with
// `post`, the NaN fill, rounding and HU rescale run on the device before the
// copy back.  For the tiled kernel the atlases are copied in chunks of whole
// atlas groups on a copy stream, and each chunk's work items run on one of
// two compute streams as soon as its copy has landed, so the copies of the
// later chunks overlap the fusion of the earlier ones.  Every item writes the same
// partial sums as in one launch, and the finalize pass adds the groups in
// their fixed order: the output is that of pct_fuse_dev bit for bit.
static int fuse_host_impl(const pct_problem* pb, const float* img, const float* atlas,
                          const float* label, float* out, int choice, bool post) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!img || !atlas || !label || !out) return PCT_ERR_INVALID;
    if (!valid_choice(choice)) return PCT_ERR_INVALID;
    const size_t N = (size_t)pb->nx * pb->ny * pb->nz;
    if (post && N >= (size_t)INT_MAX) return PCT_ERR_TOO_LARGE;   // device post chain's limit
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    const size_t NA = N * pb->n_atlas;
    HostCallState& st = host_state(dev);
    std::lock_guard<std::mutex> lk(st.mu);
    rc = prepare_host_state(st, N, NA);
    if (rc) return rc;
    hipStream_t s = st.compute, cs = st.copy;
    int sid = 0, grp = 1;
    TileGeometry g;
    const bool tiled = is_tiled_choice(pb, choice) && choose(pb, &sid, &grp, nullptr, &g);
    void* ws = nullptr;
    {
        size_t need = post ? postprocess_ws_bytes(pb->nh, pb->nx, pb->ny, pb->nz) : 0;
        if (tiled) need = std::max(need, tiled_ws_bytes(pb, g.n_groups, sid));
        if (!need) need = 256;
        rc = host_workspace(st, need, &ws);
    }
    f32x2* le = tiled && is_xterm_shape(sid)
                    ? reinterpret_cast<f32x2*>(static_cast<char*>(ws) +
                                               tiled_le_offset(pb, g.n_groups))
                    : nullptr;
    // cu:1687-1699: host->device copies
    if (!rc) rc = stage_h2d(st, st.img, img, N * 4, cs);
    if (!rc && tiled) {
        const int chunks = std::min(g.n_groups, kCopyChunks);
        for (int c = 0; c < chunks && !rc; ++c) {
            const int g0 = (int)((long long)c * g.n_groups / chunks);
            const int g1 = (int)((long long)(c + 1) * g.n_groups / chunks);
            const int a0 = g0 * g.group, a1 = std::min(pb->n_atlas, g1 * g.group);
            const size_t off = (size_t)a0 * N, len = (size_t)(a1 - a0) * N;
            rc = stage_h2d(st, st.atlas + off, atlas + off, len * 4, cs);
            if (!rc) rc = stage_h2d(st, st.label + off, label + off, len * 4, cs);
            if (rc) break;
            if (hipEventRecord(st.ev[c], cs) != hipSuccess ||
                hipStreamWaitEvent((c & 1) ? st.compute2 : s, st.ev[c], 0) != hipSuccess) {
                rc = PCT_ERR_HIP;
                break;
            }
            // the items of groups g0 .. g1-1: pointers offset to the chunk
            TiledArgs a = make_tiled_args(pb, g, st.img, st.atlas + off, st.label + off,
                                          static_cast<float*>(ws) + 2 * (size_t)g0 * N);
            a.n_atlas = a1 - a0;
            a.n_items = (int)(g.n_tiles * (g1 - g0));
            hipStream_t cstream = (c & 1) ? st.compute2 : s;
            if (le) {
                // pairs of the chunk's atlases, computed as the single launch does
                rc = launch_atlas_pairs(pb, st.atlas + off, st.label + off, a1 - a0, a.scale,
                                        le + off, cstream);
                a.le = le + off;
            }
            if (!rc) rc = launch_tiled_items(a, sid, g.hw, cstream);
        }
        if (!rc && (hipEventRecord(st.join, st.compute2) != hipSuccess ||
                    hipStreamWaitEvent(s, st.join, 0) != hipSuccess))
            rc = PCT_ERR_HIP;
        if (!rc) rc = launch_finalize(pb, g, static_cast<const float*>(ws), st.out, s);
    } else if (!rc) {
        rc = stage_h2d(st, st.atlas, atlas, NA * 4, cs);
        if (!rc) rc = stage_h2d(st, st.label, label, NA * 4, cs);
        if (!rc && (hipEventRecord(st.ev[0], cs) != hipSuccess ||
                    hipStreamWaitEvent(s, st.ev[0], 0) != hipSuccess))
            rc = PCT_ERR_HIP;
        if (!rc) rc = fuse_dev_impl(pb, st.img, st.atlas, st.label, st.out, choice, ws,
                                    st.ws_cap, false, s);
    }
    if (!rc && post) rc = postprocess_dev_ws(pb->nh, pb->nx, pb->ny, pb->nz, st.out, ws, s);
    if (!rc) rc = stage_d2h(st, out, st.out, N * 4, s);
    // every stream is drained before returning, also after an error
    if (hipStreamSynchronize(cs) != hipSuccess && !rc) rc = PCT_ERR_HIP;
    if (hipStreamSynchronize(st.compute2) != hipSuccess && !rc) rc = PCT_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = PCT_ERR_HIP;
    if (rc) (void)hipGetLastError();
    return rc;
}

namespace pct {
void release_host_buffers() {
    std::lock_guard<std::mutex> lk(g_host_mutex);
    for (auto& kv : g_host_state) {
        HostCallState* st = kv.second;
        {
            // waits for a call still running on this device; released
            // before the state (and its mutex) is deleted
            std::lock_guard<std::mutex> lk2(st->mu);
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(kv.first);
            if (st->compute) (void)hipStreamSynchronize(st->compute);
            if (st->compute2) (void)hipStreamSynchronize(st->compute2);
            if (st->copy) (void)hipStreamSynchronize(st->copy);
            free_host_buffers(*st);
            free_pinned(*st);
            for (auto& e : st->ev)
                if (e) (void)hipEventDestroy(e);
            if (st->join) (void)hipEventDestroy(st->join);
            if (st->compute) (void)hipStreamDestroy(st->compute);
            if (st->compute2) (void)hipStreamDestroy(st->compute2);
            if (st->copy) (void)hipStreamDestroy(st->copy);
            (void)hipSetDevice(cur);
        }
        delete st;
    }
    g_host_state.clear();
}
}  // namespace pct

extern "C" int pct_fuse_host(const pct_problem* pb, const float* img, const float* atlas,
                             const float* label, float* out, int choice) {
    return fuse_host_impl(pb, img, atlas, label, out, choice, false);
}

extern "C" int pct_pseudo_ct_host(const pct_problem* pb, const float* img, const float* atlas,
                                  const float* label, float* out_hu, int choice) {
    return fuse_host_impl(pb, img, atlas, label, out_hu, choice, true);
}

extern "C" int pct_tile_config_count(void) { return (int)shapes().size(); }

static void fill_config(const pct_problem* pb, int sid, int grp, double cost,
                        const TileGeometry& g, pct_tile_config* out) {
    const TileShape& s = shapes()[sid];
    out->id = sid;
    out->seg_width = s.seg_width;
    out->bx = s.bx;
    out->bz = s.bz;
    out->waves = s.waves;
    out->atlas_group = g.group;
    out->lds_bytes = g.lds_bytes;
    out->items = g.n_items;
    out->est_cost = cost;
    out->measured_ms = -1;
    out->form = s.form;
    (void)pb;
    (void)grp;
}

extern "C" int pct_tile_config_choose(const pct_problem* pb, pct_tile_config* out) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!out) return PCT_ERR_INVALID;
    int sid = 0, grp = 1;
    double cost = 0;
    TileGeometry g;
    if (!choose(pb, &sid, &grp, &cost, &g)) return PCT_ERR_UNSUPPORTED;
    fill_config(pb, sid, grp, cost, g, out);
    return PCT_OK;
}

extern "C" int pct_tile_config_force(int id, int atlas_group) {
    if (id >= (int)shapes().size()) return PCT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_cfg_mutex);
    g_forced_id = id < 0 ? -1 : id;
    g_forced_group = atlas_group;
    return PCT_OK;
}

extern "C" int pct_autotune(const pct_problem* pb, const float* d_img, const float* d_atlas,
                            const float* d_label, float* d_out, int reps, pct_stream_t stream,
                            pct_tile_config* best) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!d_img || !d_atlas || !d_label || !d_out) return PCT_ERR_INVALID;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    if ((pb->nh - 1) / 2 < 1 || (pb->nh - 1) / 2 > kMaxTiledHalfWindow) return PCT_ERR_UNSUPPORTED;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return PCT_ERR_HIP;
    double best_ms = -1;
    int best_sid = -1, best_grp = 1;
    TileGeometry best_g;
    // candidate groups: 1, the model's choice per shape, and n
    for (int sid = 0; sid < (int)shapes().size(); ++sid) {
        std::vector<int> groups = {1, pb->n_atlas};
        double bc = -1;
        int bg = 1;
        for (int grp = 1; grp <= pb->n_atlas; ++grp) {
            double c = model_cost(pb, sid, grp, nullptr);
            if (c >= 0 && (bc < 0 || c < bc * 0.999)) {
                bc = c;
                bg = grp;
            }
        }
        if (bc < 0) continue;
        groups.push_back(bg);
        std::sort(groups.begin(), groups.end());
        groups.erase(std::unique(groups.begin(), groups.end()), groups.end());
        for (int grp : groups) {
            TileGeometry g;
            if (!make_geometry(pb, sid, grp, &g)) continue;
            void* ws = nullptr;
            const size_t need = tiled_ws_bytes(pb, g.n_groups, sid);
            rc = ensure_workspace(dev, need, &ws);
            if (rc) break;
            // warm-up
            rc = launch_tiled(pb, d_img, d_atlas, d_label, d_out, ws, need, sid, g, s);
            if (rc) break;
            (void)hipEventRecord(e0, s);
            for (int i = 0; i < reps && !rc; ++i)
                rc = launch_tiled(pb, d_img, d_atlas, d_label, d_out, ws, need, sid, g, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = PCT_ERR_HIP;
            if (rc) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= reps;
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best_sid = sid;
                best_grp = grp;
                best_g = g;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_sid < 0) return PCT_ERR_UNSUPPORTED;
    {
        std::lock_guard<std::mutex> lk(g_cfg_mutex);
        g_tuned[ShapeKey{pb->nx, pb->ny, pb->nz, pb->n_atlas, pb->nh}] = {best_sid, best_grp};
    }
    if (best) {
        fill_config(pb, best_sid, best_grp, model_cost(pb, best_sid, best_grp, nullptr), best_g,
                    best);
        best->measured_ms = best_ms;
    }
    return PCT_OK;
}

extern "C" void pct_release(void) {
    release_batch_graphs();
    release_learned_models();
    release_host_buffers();
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    for (auto& kv : g_dev) {
        if (kv.second.ws) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(kv.first);
            (void)hipDeviceSynchronize();
            (void)hipFree(kv.second.ws);
            (void)hipSetDevice(cur);
            kv.second.ws = nullptr;
            kv.second.ws_bytes = 0;
        }
    }
}
