// Device-side building blocks of the tiled fusion kernels, shared by the
// single-problem kernels (pct_fuse.hip) and the persistent batched kernel
// (pct_batch.hip).  See pct_fuse.hip for the design.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pct_internal.h"

namespace pct {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// lane l receives lane l-1 (lane 0 receives 0).
__device__ __forceinline__ float lane_prev(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
// lane l receives lane l+1 (lane 63 receives 0).
__device__ __forceinline__ float lane_next(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

// Work items are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of items so that tiles that share halos share an L2.
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int per = n >> 3, rem = n & 7;
    const int xcd = b & 7, local = b >> 3;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// The reference's rule (cpp:1437-1443) in float: DBL_EPSILON = 2^-52 is exact.
__device__ __forceinline__ float fused_value(float P, float S) {
    const float eps = 2.220446049250313080847e-16f;
    return (fabsf(P) > eps && fabsf(S) > eps) ? (P / S) : __builtin_nanf("");
}

// ---------------------------------------------------------------------------
// tiled kernel
// ---------------------------------------------------------------------------

// Native two-float vector: its loads and stores stay single 64-bit accesses
// (HIP's float2 struct is copied member by member).
typedef float f32x2 __attribute__((ext_vector_type(2)));

struct TiledArgs {
    const float* img;
    const float* atlas;
    const float* label;
    float* ws;        // [n_groups][2][N]
    long long N;      // nx*ny*nz
    int nx, ny, nz;
    int r;            // hw + 1
    int tiles_y, tiles_x;
    int n_tiles;
    int n_items;
    int n_atlas, group;
    float scale;      // sqrt(log2(e)/div)
    // optional item filter: when set, items whose byte is 0 are skipped
    const unsigned char* only;
    // cross-term shapes only: per atlas voxel the pair (L * E, E) with
    // E = 2^-AA and AA the 3x3x3 sum of the squared scaled atlas MRI
    // ([n_atlas][N] pairs, written by atlas_sq_box)
    const f32x2* le;
};

template <int W, int BX, int BZ, int NW, int HW>
struct TiledCfg {
    static constexpr int BXv = BX, BZv = BZ, HWv = HW, NWv = NW;
    static constexpr int S = 64 / W;           // segments per wave
    static constexpr int NSEG = NW * S;        // segments per workgroup (stacked in x)
    static constexpr int WX = NSEG * BX;       // output rows per workgroup
    static constexpr int OY = W - 2;           // outputs per segment in y
    static constexpr int CW = W + 2 * HW;      // LDS row length (y)
    static constexpr int RX = WX + 2 + 2 * HW; // LDS rows per plane (x)
    static constexpr int PS = RX * CW;         // LDS plane stride
    static constexpr int APLANES = BZ + 2;     // atlas MRI planes per stage
    static constexpr int LPLANES = BZ;         // atlas CT planes per stage
    static constexpr int LOFF = APLANES * PS;  // CT region offset
    static constexpr int LDS_FLOATS = (APLANES + LPLANES) * PS;
    static constexpr int LDS_BYTES = LDS_FLOATS * 4;
    static constexpr int COLCHUNKS = (CW + 63) / 64;
    // ds_read_b32 takes a 16-bit byte offset: planes are addressed from
    // several base registers, PPB planes (plus row offsets) per base.
    // The immediate also carries the candidate's dy (up to 2 HW).
    static constexpr int PPB = (16383 - (BX + 1) * CW - 2 * HW) / PS + 1;
    static_assert(PPB >= 1, "one plane must fit an LDS immediate window");
    static constexpr int NBA = (APLANES + PPB - 1) / PPB;   // base registers, MRI planes
    static constexpr int NBL = (LPLANES + PPB - 1) / PPB;   // base registers, CT planes
    // the same for the dx sweep, whose immediates carry the row (up to
    // 2 HW + BX + 1) and no dy
    static constexpr int PPBS = (16383 - (2 * HW + BX + 1) * CW) / PS + 1;
    static_assert(PPBS >= 1, "one plane must fit an LDS immediate window");
    static constexpr int NBAS = (APLANES + PPBS - 1) / PPBS;
    static constexpr int NBLS = (LPLANES + PPBS - 1) / PPBS;
    static_assert(64 % W == 0, "segment width must divide the wavefront");
    static_assert(BX % 2 == 0 && BZ % 2 == 0, "pairwise box sums need even blocks");
};

// Hides an LDS element index from the optimiser so that it stays one base
// register instead of being folded into offsets that do not fit an LDS
// immediate.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// All 2 HW + 1 dx candidates of one dy with squared differences.  `base` is
// the LDS element index of plane 0, row 0 of the lane's window for dx = 0.
// Candidate dx reads MRI rows dx .. dx + BX + 1, so consecutive candidates
// share BX + 1 of their BX + 2 rows: the rows stay in registers and each
// candidate after the first reads one new row (BZ + 2 reads instead of
// (BZ + 2)(BX + 2)); CT rows are shared the same way.  The planes are
// addressed from a few base registers hidden from the optimiser (`opaque`),
// so each stays one register and the plane and row offsets ride in the
// ds_read immediates.  Per candidate the lane squares its (BZ+2) x (BX+2)
// differences, sums them pairwise over x and then z in registers and over y
// across lanes (two DPP adds), and weighs with w = 2^-SSD' on MRI values
// pre-scaled by sqrt(log2(e) / div).
template <class C, int BX, int BZ, int HW>
__device__ __forceinline__ void tiled_dx_sweep(const float* __restrict__ lds, int base,
                                               const float (&tv)[BZ + 2][BX + 2],
                                               float (&P)[BZ][BX], float (&S)[BZ][BX]) {
    int ba[C::NBAS], bl[C::NBLS];
#pragma unroll
    for (int b = 0; b < C::NBAS; ++b) ba[b] = opaque(base + b * C::PPBS * C::PS);
#pragma unroll
    for (int b = 0; b < C::NBLS; ++b) bl[b] = opaque(base + C::LOFF + b * C::PPBS * C::PS);
    constexpr int NR = BX + 2 + 2 * HW;
    float rows[NR][BZ + 2];
    auto load_row = [&](int r) {
#pragma unroll
        for (int pz = 0; pz < BZ + 2; ++pz)
            rows[r][pz] = lds[ba[pz / C::PPBS] + (pz % C::PPBS) * C::PS + r * C::CW];
    };
#pragma unroll
    for (int r = 0; r < BX + 1; ++r) load_row(r);
#pragma unroll
    for (int dx = 0; dx <= 2 * HW; ++dx) {
        load_row(dx + BX + 1);
        float ex[BZ + 2][BX];
#pragma unroll
        for (int pz = 0; pz < BZ + 2; ++pz) {
            float d[BX + 2];
#pragma unroll
            for (int i = 0; i < BX + 2; ++i) {
                const float df = tv[pz][i] - rows[dx + i][pz];
                d[i] = df * df;
            }
#pragma unroll
            for (int i = 0; i < BX; i += 2) {
                const float q = d[i + 1] + d[i + 2];
                ex[pz][i] = d[i] + q;
                ex[pz][i + 1] = q + d[i + 3];
            }
        }
#pragma unroll
        for (int j = 0; j < BZ; j += 2) {
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                const float q = ex[j + 1][i] + ex[j + 2][i];
                const float e0 = ex[j][i] + q;
                const float e1 = q + ex[j + 3][i];
                const float s0 = lane_next(e0) + (lane_prev(e0) + e0);
                const float s1 = lane_next(e1) + (lane_prev(e1) + e1);
                const float w0 = __builtin_amdgcn_exp2f(-s0);
                const float w1 = __builtin_amdgcn_exp2f(-s1);
                const float l0 =
                    lds[bl[j / C::PPBS] + (j % C::PPBS) * C::PS + (dx + i + 1) * C::CW];
                const float l1 = lds[bl[(j + 1) / C::PPBS] + ((j + 1) % C::PPBS) * C::PS +
                                     (dx + i + 1) * C::CW];
                P[j][i] = fmaf(w0, l0, P[j][i]);
                S[j][i] += w0;
                P[j + 1][i] = fmaf(w1, l1, P[j + 1][i]);
                S[j + 1][i] += w1;
            }
        }
    }
}

// Staging loads read clamped positions (always inside the volume) and
// replace the values of positions outside it after the load.  A load under
// a branch would be waited for before the branch closes, one load at a
// time; without one, a lane's loads of a plane batch are all in flight
// together.
__device__ __forceinline__ int clampi(int v, int n) { return min(max(v, 0), n - 1); }

// The lane's slab columns c = lane + 64 cc: clamped volume y, and whether
// the column is a slab column inside the volume.
template <class C>
__device__ __forceinline__ void lane_columns(int ys, int ny, int lane, int (&ycl)[C::COLCHUNKS],
                                             bool (&yin)[C::COLCHUNKS]) {
#pragma unroll
    for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
        const int c = lane + cc * 64;
        const int yy = ys + c;
        yin[cc] = c < C::CW && yy >= 0 && yy < ny;
        ycl[cc] = clampi(yy, ny);
    }
}

// Copies the slab of one (atlas, dz) stage into LDS: plane p < APLANES holds
// the atlas MRI at z = zdz-1+p (scaled), plane APLANES+j the atlas CT at
// z = zdz+j; row rr is volume x = xs+rr, column c is volume y = ys+c; zero
// outside the volume.  Each wave issues the loads of all its rows of two
// planes before it writes any of them, so global latency is paid once per
// pair of planes.
template <class C, int NW>
__device__ __forceinline__ void stage_slab(float* __restrict__ lds, const float* __restrict__ A,
                                           const float* __restrict__ L, float scale, int zdz,
                                           int xs, int ys, int nx, int ny, int nz,
                                           long long plane, int wv, int lane) {
    constexpr int NP = C::APLANES + C::LPLANES;
    constexpr int RPW = (C::RX + NW - 1) / NW;   // rows per wave and plane
    constexpr int PB = 2;                         // planes per batch
    static_assert(NP % PB == 0, "plane batches must tile the slab");
    int ycl[C::COLCHUNKS];
    bool yin[C::COLCHUNKS];
    lane_columns<C>(ys, ny, lane, ycl, yin);
#pragma unroll 1
    for (int p0 = 0; p0 < NP; p0 += PB) {
        float v[PB][RPW][C::COLCHUNKS];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int p = p0 + q;
            const bool is_mri = p < C::APLANES;
            const int zz = is_mri ? zdz - 1 + p : zdz + (p - C::APLANES);
            const float* __restrict__ src = is_mri ? A : L;
            const float mul = is_mri ? scale : 1.f;
            const bool zok = zz >= 0 && zz < nz;
            const float* __restrict__ sz = src + (long long)clampi(zz, nz) * plane;
#pragma unroll
            for (int k = 0; k < RPW; ++k) {
                const int rr = wv + k * NW;
                const int xx = xs + rr;
                const bool xok = zok && rr < C::RX && xx >= 0 && xx < nx;
                const float* __restrict__ sr = sz + (long long)clampi(xx, nx) * ny;
#pragma unroll
                for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                    const float raw = sr[ycl[cc]];
                    v[q][k][cc] = (xok && yin[cc]) ? raw * mul : 0.f;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < PB; ++q)
#pragma unroll
            for (int k = 0; k < RPW; ++k) {
                const int rr = wv + k * NW;
#pragma unroll
                for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                    const int c = lane + cc * 64;
                    if (rr < C::RX && c < C::CW) lds[(p0 + q) * C::PS + rr * C::CW + c] = v[q][k][cc];
                }
            }
    }
}

// Zeros read by lanes whose slab position lies outside the volume.
__device__ const float kZeros[64] = {};

// The slab of stage_slab, copied with direct global-to-LDS loads
// (global_load_lds_dword): every lane of a wave instruction fills one of 64
// consecutive LDS floats of a plane, so a plane is PS/64 instructions, all of
// the slab's loads are in flight at once, and no VGPR holds the data.  Wave w
// copies chunks j = w, w + NW, ... of every plane, so a lane's in-plane
// source offsets depend only on the tile: DirectPlan computes them once per
// item, and a copy costs one address add per instruction.
template <class C, int NW>
struct DirectPlan {
    static constexpr int CH = (C::PS + 63) / 64;      // 64-float chunks per plane
    static constexpr int JN = (CH + NW - 1) / NW;     // chunks per wave and plane
    // x * ny + y of the lane's source voxel within a plane, -1 outside the
    // volume (reads zeros), -2 past the end of the plane (no copy)
    int off[JN];

    __device__ __forceinline__ void init(int xs, int ys, int nx, int ny, int wv, int lane) {
#pragma unroll
        for (int k = 0; k < JN; ++k) {
            const int j = wv + k * NW;
            const int p = j * 64 + lane;
            const int row = p / C::CW, col = p - (p / C::CW) * C::CW;
            const int xx = xs + row, yy = ys + col;
            int o = -2;
            if (j < CH && p < C::PS)
                o = (xx >= 0 && xx < nx && yy >= 0 && yy < ny) ? xx * ny + yy : -1;
            off[k] = o;
        }
    }
};

// Copies the slab of stage_slab from the unscaled atlas MRI; positions
// outside the volume read zeros.  Once its own loads have landed, each lane
// multiplies the MRI positions they wrote by `scale`: the product stage_slab
// forms before its LDS store (zeros stay zero), so both stagings hold the same
// bits.  Ends with a workgroup barrier.
template <class C, int NW>
__device__ __forceinline__ void stage_slab_direct(float* __restrict__ lds,
                                                  const float* __restrict__ A,
                                                  const float* __restrict__ L, float scale,
                                                  int zdz, int nz, long long plane, int wv,
                                                  const DirectPlan<C, NW>& dp) {
    typedef __attribute__((address_space(3))) void* lds_ptr;
    constexpr int NP = C::APLANES + C::LPLANES;
    using P = DirectPlan<C, NW>;
#pragma unroll 2
    for (int pl = 0; pl < NP; ++pl) {
        const bool is_mri = pl < C::APLANES;
        const int zz = is_mri ? zdz - 1 + pl : zdz + (pl - C::APLANES);
        const bool zok = zz >= 0 && zz < nz;
        const float* base = (is_mri ? A : L) + (zok ? zz : 0) * plane;
#pragma unroll
        for (int k = 0; k < P::JN; ++k) {
            const int j = wv + k * NW;
            const int o = dp.off[k];
            if (o >= -1) {
                const float* src = (zok && o >= 0) ? base + o : kZeros;
                __builtin_amdgcn_global_load_lds(src, (lds_ptr)(lds + pl * C::PS + j * 64), 4,
                                                 0, 0);
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __asm__ volatile("" ::: "memory");
    // lane l of load (plane pl, chunk j) wrote LDS element pl * PS + j * 64 + l
    const int lane = (int)(threadIdx.x & 63u);
#pragma unroll
    for (int pl = 0; pl < C::APLANES; ++pl)
#pragma unroll
        for (int k = 0; k < P::JN; ++k)
            if (dp.off[k] >= -1) {
                float* q = lds + pl * C::PS + (wv + k * NW) * 64 + lane;
                *q = *q * scale;
            }
    __syncthreads();
}

// Scaled target values of a lane's BZ x BX block plus a one-voxel halo in z
// and x, zero outside the volume.  The loads use clamped (always valid)
// addresses and are selected afterwards, so they are all issued before the
// first wait.
template <int BX, int BZ>
__device__ __forceinline__ void load_target_block(const float* __restrict__ img, float scale,
                                                  int z0, int xb, int y, int nx, int ny, int nz,
                                                  long long plane,
                                                  float (&tv)[BZ + 2][BX + 2]) {
    const int yc = min(y, ny - 1);
#pragma unroll
    for (int pz = 0; pz < BZ + 2; ++pz)
#pragma unroll
        for (int i = 0; i < BX + 2; ++i) {
            const int zz = z0 - 1 + pz, xx = xb - 1 + i;
            const float v = img[min(zz, nz - 1) * plane + (long long)min(xx, nx - 1) * ny + yc];
            tv[pz][i] = (zz < nz && xx < nx && y < ny) ? v * scale : 0.f;
        }
}

// Whether wave wv of a workgroup has an output row before x_end.  A tile
// column that runs past the interior (x_end = nx - r) holds waves whose rows
// all lie beyond it: they stage their share of every slab and take part in
// the barriers, but skip the candidates, whose sums would never be stored.
// Uniform over the wave (read from the first lane), so the skip is a scalar
// branch.
template <class C, int BX>
__device__ __forceinline__ bool wave_has_rows(int x0, int x_end, int wv) {
    return x0 + __builtin_amdgcn_readfirstlane(wv) * C::S * BX < x_end;
}

// One work item (tile, atlas group) of one problem: the whole workgroup runs
// it.  Shared by the single-problem kernel and the persistent batched kernel,
// so a problem gives bit-identical results through either.
template <int W, int BX, int BZ, int NW, int HW>
__device__ __forceinline__ void tiled_item(const TiledArgs& args, int item,
                                           float* __restrict__ lds) {
    using C = TiledCfg<W, BX, BZ, NW, HW>;
    const int g = item / args.n_tiles;
    int t = item - g * args.n_tiles;
    const int ty = t % args.tiles_y;
    t /= args.tiles_y;
    const int tx = t % args.tiles_x;
    const int tz = t / args.tiles_x;

    const int tid = threadIdx.x;
    const int wv = tid >> 6;
    const int lane = tid & 63;
    const int seg = lane / W;
    const int ly = lane % W;
    const int sg = wv * C::S + seg;

    const int nx = args.nx, ny = args.ny, nz = args.nz, r = args.r;
    const long long plane = (long long)nx * ny;
    const int y0 = r + ty * C::OY;
    const int x0 = r + tx * C::WX;
    const int z0 = r + tz * BZ;
    const int y = y0 - 1 + ly;        // >= 0 because r >= 1
    const int xb = x0 + sg * BX;      // first output row of this lane

    // Scaled target values of the lane's block plus a one-voxel halo in z
    // and x (y halo comes from the neighbouring lanes).  Zero outside the
    // volume: such values only reach outputs that are never stored.
    float tv[BZ + 2][BX + 2];
    load_target_block<BX, BZ>(args.img, args.scale, z0, xb, y, nx, ny, nz, plane, tv);

    float P[BZ][BX], S[BZ][BX];
#pragma unroll
    for (int j = 0; j < BZ; ++j)
#pragma unroll
        for (int i = 0; i < BX; ++i) P[j][i] = S[j][i] = 0.f;

    const int lane_base = sg * BX * C::CW + ly;
    const int xs = x0 - 1 - HW;   // volume x of LDS row 0
    const int ys = y0 - 1 - HW;   // volume y of LDS column 0
    const int a_begin = g * args.group;
    const int a_end = min(args.n_atlas, a_begin + args.group);
    const bool live = wave_has_rows<C, BX>(x0, nx - r, wv);

    for (int a = a_begin; a < a_end; ++a) {
        const float* __restrict__ A = args.atlas + (long long)a * args.N;
        const float* __restrict__ L = args.label + (long long)a * args.N;
        for (int dz = -HW; dz <= HW; ++dz) {
            __syncthreads();
            stage_slab<C, NW>(lds, A, L, args.scale, z0 + dz, xs, ys, nx, ny, nz, plane, wv, lane);
            __syncthreads();
            if (!live) continue;
            // dy rolled, dx unrolled: consecutive dx share MRI and CT rows
#pragma unroll 1
            for (int dy = 0; dy <= 2 * HW; ++dy)
                tiled_dx_sweep<C, BX, BZ, HW>(lds, lane_base + dy, tv, P, S);
        }
    }

    // Store the group's partial sums for the lane's valid outputs.
    if (ly >= 1 && ly <= W - 2 && y < ny - r) {
        float* __restrict__ wsP = args.ws + (long long)(2 * g) * args.N;
        float* __restrict__ wsS = wsP + args.N;
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                const int z = z0 + j, x = xb + i;
                if (z < nz - r && x < nx - r) {
                    const long long idx = z * plane + (long long)x * ny + y;
                    wsP[idx] = P[j][i];
                    wsS[idx] = S[j][i];
                }
            }
    }
}

template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) void fuse_tiled(TiledArgs args) {
    extern __shared__ float lds[];
    const int item = xcd_contiguous(blockIdx.x, args.n_items);
    if (args.only && !args.only[item]) return;   // uniform over the workgroup
    tiled_item<W, BX, BZ, NW, HW>(args, item, lds);
}

// ---------------------------------------------------------------------------
// tiled kernel, cross-term form
// ---------------------------------------------------------------------------
//
// With T' and A' the scaled target and atlas MRI, the patch SSD of output v
// and candidate d expands as
//     SSD' = TT(v) + AA(v+d) - C(v, d),
//     TT = sum_p T'(v+p)^2,  AA = sum_p A'(v+d+p)^2,  C = sum_p 2 T'(v+p) A'(v+d+p),
// so the weight is w = 2^-SSD' = 2^-TT(v) * 2^C * 2^-AA(v+d).  2^-TT(v) is
// one factor per output: it is applied to P and S once, after the last
// candidate.  2^-AA(v+d) depends only on the atlas position: AA comes from
// one pass over the atlases per call (atlas_sq_box), and each slab carries
// the pair (L~, E) = (L * 2^-AA, 2^-AA) per position in place of the CT.
// Per candidate the lane forms C with the same separable box sums as the
// squared-difference kernel, its x sums now FMA chains of products (one
// instruction per voxel instead of a difference and a square), then
// w~ = 2^C, P += w~ L~, S += w~ E.
//
// Every LDS read is 64 bits wide, which costs one instruction slot like a
// 32-bit read on gfx950: the MRI planes are interleaved in pairs, so a read
// returns two planes at a position, and (L~, E) is one pair.  The candidates
// of one dy are swept over dx with the MRI rows kept in registers (one new
// row per candidate), and all their reads address one base register per dy
// (plane and row offsets ride in the 16-bit immediates).
//
// The expanded form loses absolute precision in proportion to TT + AA, and
// 2^C must not overflow, so it is used only where every |T'| and |A'| is at
// most kXtBound: then C <= 2 * 27 * kXtBound^2 = 3.4 and the exponent error
// stays within a few float ulps of 3.4.  An item whose target values exceed
// the bound runs the squared-difference form throughout; a slab whose atlas
// values (or CT values, beyond kXtLabelBound) exceed it is weighed with the
// squared differences and w = 2^(TT - SSD'), the same scaled weight.

constexpr float kXtBound = 0.25f;
constexpr float kXtLabelBound = 1.8446744e19f;   // 2^64: bounds P / 2^-TT

typedef __attribute__((address_space(3))) const f32x2 lds_pair_t;

// LDS byte address of an element of the dynamic shared array.
__device__ __forceinline__ unsigned lds_byte_address(const float* p) {
    return (unsigned)(size_t)(__attribute__((address_space(3))) const float*)p;
}

__device__ __forceinline__ unsigned opaque_u(unsigned v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Slab of the cross-term form: NPP planes of float pairs, PS pairs each.
// Pair planes 0 .. BZ/2 hold the MRI planes (2k, 2k+1) interleaved; pair
// planes BZ/2 + 1 + j hold (L~, E) of CT plane j.
template <class C>
struct XtCfg {
    static_assert(C::APLANES % 2 == 0, "MRI planes are interleaved in pairs");
    static constexpr int MPAIRS = C::APLANES / 2;
    static constexpr int NPP = MPAIRS + C::LPLANES;
    static constexpr int LDS_FLOATS = 2 * NPP * C::PS;
    static constexpr int LDS_BYTES = LDS_FLOATS * 4;
    static_assert(LDS_FLOATS >= C::LDS_FLOATS, "the squared-difference slab must fit too");
    // pair planes per base register: the 16-bit byte immediate of a 64-bit
    // read also carries the row (up to BX + 1) and dy (up to 2 HW)
    static constexpr int PPP = (8191 - ((C::BXv + 1) * C::CW + 2 * C::HWv)) / C::PS + 1;
    static_assert(PPP >= 1, "one pair plane must fit an LDS immediate window");
    static constexpr int NB = (NPP + PPP - 1) / PPP;
    // the same for the dx sweep, whose immediates carry the row (up to
    // 2 HW + BX + 1) and no dy
    static constexpr int PPPS = (8191 - (2 * C::HWv + C::BXv + 1) * C::CW) / C::PS + 1;
    static_assert(PPPS >= 1, "one pair plane must fit an LDS immediate window");
    static constexpr int NBS = (NPP + PPPS - 1) / PPPS;
    // float index of MRI plane pz, and of the (L~, E) pair of CT plane j, at
    // in-plane position p
    static __device__ __forceinline__ int mri(int pz, int p) {
        return 2 * ((pz >> 1) * C::PS + p) + (pz & 1);
    }
    static __device__ __forceinline__ int le(int j, int p) { return 2 * ((MPAIRS + j) * C::PS + p); }
};

// Byte-address base registers of one candidate row: in-plane position `pos`
// (row, column of the lane's window for dy = 0) in pair planes 0, PPP, ...
template <class C>
__device__ __forceinline__ void make_bases_xt(unsigned lds0, int pos,
                                              unsigned (&b)[XtCfg<C>::NB]) {
    using X = XtCfg<C>;
#pragma unroll
    for (int k = 0; k < X::NB; ++k) b[k] = opaque_u(lds0 + 8u * (unsigned)(pos + k * X::PPP * C::PS));
}

// The pair at pair plane `pl`, in-plane offset `off` from the bases.
template <class C>
__device__ __forceinline__ f32x2 pair_at(const unsigned (&b)[XtCfg<C>::NB], int pl, int off) {
    using X = XtCfg<C>;
    const lds_pair_t* q = (const lds_pair_t*)(size_t)b[pl / X::PPP];
    return q[(pl % X::PPP) * C::PS + off];
}

// Bases and pair reads of the dx sweep (see XtCfg::PPPS).
template <class C>
__device__ __forceinline__ void make_bases_sweep(unsigned lds0, int pos,
                                                 unsigned (&b)[XtCfg<C>::NBS]) {
    using X = XtCfg<C>;
#pragma unroll
    for (int k = 0; k < X::NBS; ++k)
        b[k] = opaque_u(lds0 + 8u * (unsigned)(pos + k * X::PPPS * C::PS));
}

template <class C>
__device__ __forceinline__ f32x2 pair_at_sweep(const unsigned (&b)[XtCfg<C>::NBS], int pl,
                                               int off) {
    using X = XtCfg<C>;
    const lds_pair_t* q = (const lds_pair_t*)(size_t)b[pl / X::PPPS];
    return q[(pl % X::PPPS) * C::PS + off];
}

// The lane's MRI values of one candidate: a[pz][i] = plane pz, row i.
template <class C, int BX, int BZ>
__device__ __forceinline__ void load_mri_xt(const unsigned (&b)[XtCfg<C>::NB],
                                            float (&a)[BZ + 2][BX + 2]) {
    using X = XtCfg<C>;
#pragma unroll
    for (int k = 0; k < X::MPAIRS; ++k)
#pragma unroll
        for (int i = 0; i < BX + 2; ++i) {
            const f32x2 v = pair_at<C>(b, k, i * C::CW);
            a[2 * k][i] = v.x;
            a[2 * k + 1][i] = v.y;
        }
}

// All 2 HW + 1 dx candidates of one dy in cross-term form.  Candidate dx
// reads MRI rows dx .. dx + BX + 1, so consecutive candidates share BX + 1
// of their BX + 2 rows: the rows stay in registers and each candidate after
// the first reads one new row (MPAIRS 64-bit reads instead of
// MPAIRS (BX + 2)).  Row and plane offsets ride in the immediates of the
// dy's bases.  Per candidate the lane forms C with separable sums (x sums as
// FMA chains of products, then z pairwise, then y across lanes), then
// w~ = 2^C, P += w~ L~, S += w~ E; t2 holds 2 T'.
template <class C, int BX, int BZ, int HW>
__device__ __forceinline__ void xterm_dx_sweep(const unsigned (&b)[XtCfg<C>::NBS],
                                               const float (&t2)[BZ + 2][BX + 2],
                                               float (&P)[BZ][BX], float (&S)[BZ][BX]) {
    using X = XtCfg<C>;
    constexpr int NR = BX + 2 + 2 * HW;
    float rows[NR][BZ + 2];
    auto load_row = [&](int r) {
#pragma unroll
        for (int k = 0; k < X::MPAIRS; ++k) {
            const f32x2 v = pair_at_sweep<C>(b, k, r * C::CW);
            rows[r][2 * k] = v.x;
            rows[r][2 * k + 1] = v.y;
        }
    };
#pragma unroll
    for (int r = 0; r < BX + 1; ++r) load_row(r);
#pragma unroll
    for (int dx = 0; dx <= 2 * HW; ++dx) {
        load_row(dx + BX + 1);
        float ex[BZ + 2][BX];
#pragma unroll
        for (int pz = 0; pz < BZ + 2; ++pz) {
#pragma unroll
            for (int i = 0; i < BX; i += 2) {
                float q = t2[pz][i + 1] * rows[dx + i + 1][pz];
                q = fmaf(t2[pz][i + 2], rows[dx + i + 2][pz], q);
                ex[pz][i] = fmaf(t2[pz][i], rows[dx + i][pz], q);
                ex[pz][i + 1] = fmaf(t2[pz][i + 3], rows[dx + i + 3][pz], q);
            }
        }
#pragma unroll
        for (int j = 0; j < BZ; j += 2) {
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                const float q = ex[j + 1][i] + ex[j + 2][i];
                const float e0 = ex[j][i] + q;
                const float e1 = q + ex[j + 3][i];
                const float c0 = lane_next(e0) + (lane_prev(e0) + e0);
                const float c1 = lane_next(e1) + (lane_prev(e1) + e1);
                const float w0 = __builtin_amdgcn_exp2f(c0);
                const float w1 = __builtin_amdgcn_exp2f(c1);
                const f32x2 g0 = pair_at_sweep<C>(b, X::MPAIRS + j, (dx + i + 1) * C::CW);
                const f32x2 g1 = pair_at_sweep<C>(b, X::MPAIRS + j + 1, (dx + i + 1) * C::CW);
                P[j][i] = fmaf(w0, g0.x, P[j][i]);
                S[j][i] = fmaf(w0, g0.y, S[j][i]);
                P[j + 1][i] = fmaf(w1, g1.x, P[j + 1][i]);
                S[j + 1][i] = fmaf(w1, g1.y, S[j + 1][i]);
            }
        }
    }
}

// One candidate of a slab the cross-term form cannot take (restaged with
// the raw CT in the first float of each pair): squared differences
// (T' = t2 / 2, exact), weight 2^(TT - SSD') on the scale of the cross-term
// accumulators.
template <class C, int BX, int BZ>
__device__ __forceinline__ void scaled_ssd_step(const unsigned (&b)[XtCfg<C>::NB],
                                                const float (&t2)[BZ + 2][BX + 2],
                                                const float (&tt)[BZ][BX], float (&P)[BZ][BX],
                                                float (&S)[BZ][BX]) {
    using X = XtCfg<C>;
    float a[BZ + 2][BX + 2];
    load_mri_xt<C, BX, BZ>(b, a);
    float ex[BZ + 2][BX];
#pragma unroll
    for (int pz = 0; pz < BZ + 2; ++pz) {
        float d[BX + 2];
#pragma unroll
        for (int i = 0; i < BX + 2; ++i) {
            const float df = fmaf(0.5f, t2[pz][i], -a[pz][i]);
            d[i] = df * df;
        }
#pragma unroll
        for (int i = 0; i < BX; i += 2) {
            const float q = d[i + 1] + d[i + 2];
            ex[pz][i] = d[i] + q;
            ex[pz][i + 1] = q + d[i + 3];
        }
    }
#pragma unroll
    for (int j = 0; j < BZ; j += 2) {
#pragma unroll
        for (int i = 0; i < BX; ++i) {
            const float q = ex[j + 1][i] + ex[j + 2][i];
            const float e0 = ex[j][i] + q;
            const float e1 = q + ex[j + 3][i];
            const float s0 = lane_next(e0) + (lane_prev(e0) + e0);
            const float s1 = lane_next(e1) + (lane_prev(e1) + e1);
            const float w0 = __builtin_amdgcn_exp2f(tt[j][i] - s0);
            const float w1 = __builtin_amdgcn_exp2f(tt[j + 1][i] - s1);
            const float l0 = pair_at<C>(b, X::MPAIRS + j, (i + 1) * C::CW).x;
            const float l1 = pair_at<C>(b, X::MPAIRS + j + 1, (i + 1) * C::CW).x;
            P[j][i] = fmaf(w0, l0, P[j][i]);
            S[j][i] += w0;
            P[j + 1][i] = fmaf(w1, l1, P[j + 1][i]);
            S[j + 1][i] += w1;
        }
    }
}

// Copies the slab of one (atlas, dz) stage in the cross-term layout: MRI
// planes z = zdz-1 .. zdz+BZ (scaled, interleaved in pairs) and, for CT
// planes z = zdz .. zdz+BZ-1, the pairs (L * E, E) of the AA pre-pass;
// positions outside the volume hold zeros.  Returns whether any lane of the
// wave staged an MRI value beyond kXtBound (or not a number) or a CT value
// beyond kXtLabelBound (the pre-pass writes +inf in place of L * E for a CT
// value beyond the bound or not a number, and |L * E| <= |L| otherwise).
// The checks are wave masks: one compare per value, the OR on the scalar
// unit.
template <class C, int NW>
__device__ __forceinline__ int stage_slab_xt(float* __restrict__ lds, const float* __restrict__ A,
                                             const f32x2* __restrict__ LE, float scale, int zdz,
                                             int xs, int ys, int nx, int ny, int nz,
                                             long long plane, int wv, int lane) {
    using X = XtCfg<C>;
    constexpr int RPW = (C::RX + NW - 1) / NW;   // rows per wave and plane
    unsigned long long bad = 0;
    f32x2* __restrict__ lp = reinterpret_cast<f32x2*>(lds);
    int ycl[C::COLCHUNKS];
    bool yin[C::COLCHUNKS];
    lane_columns<C>(ys, ny, lane, ycl, yin);
    // MRI: one pair plane (two volume planes) per round
#pragma unroll 1
    for (int k = 0; k < X::MPAIRS; ++k) {
        f32x2 v[RPW][C::COLCHUNKS];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int zz = zdz - 1 + 2 * k + h;
            const bool zok = zz >= 0 && zz < nz;
            const float* __restrict__ Az = A + (long long)clampi(zz, nz) * plane;
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int rr = wv + r * NW;
                const int xx = xs + rr;
                const bool xok = zok && rr < C::RX && xx >= 0 && xx < nx;
                const float* __restrict__ Ar = Az + (long long)clampi(xx, nx) * ny;
#pragma unroll
                for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                    const float raw = Ar[ycl[cc]];
                    const float t = (xok && yin[cc]) ? raw * scale : 0.f;
                    bad |= __ballot(!(fabsf(t) <= kXtBound));
                    v[r][cc][h] = t;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rr = wv + r * NW;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                const int c = lane + cc * 64;
                if (rr < C::RX && c < C::CW) lp[k * C::PS + rr * C::CW + c] = v[r][cc];
            }
        }
    }
    // CT planes: the (L * E, E) pairs
#pragma unroll 1
    for (int j = 0; j < C::LPLANES; ++j) {
        const int zz = zdz + j;
        const bool zok = zz >= 0 && zz < nz;
        const long long zoff = (long long)clampi(zz, nz) * plane;
        f32x2 g[RPW][C::COLCHUNKS];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int xx = xs + wv + r * NW;
            const long long ro = zoff + (long long)clampi(xx, nx) * ny;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) g[r][cc] = LE[ro + ycl[cc]];
        }
        // every load is issued before the first select: without the barrier
        // the compiler moves each 64-bit load under its own branch and waits
        // for it there
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) asm volatile("" : "+v"(g[r][cc]));
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rr = wv + r * NW;
            const int xx = xs + rr;
            const bool xok = zok && rr < C::RX && xx >= 0 && xx < nx;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc)
                g[r][cc] = (xok && yin[cc]) ? g[r][cc] : f32x2{0.f, 0.f};
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rr = wv + r * NW;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                const int c = lane + cc * 64;
                bad |= __ballot(fabsf(g[r][cc].x) > kXtLabelBound);
                if (rr < C::RX && c < C::CW) lp[(X::MPAIRS + j) * C::PS + rr * C::CW + c] = g[r][cc];
            }
        }
    }
    return bad != 0;
}

// Rewrites the first float of every (L~, E) pair of the slab with the raw CT
// (zero outside the volume), for the squared-difference weighing of a slab
// out of the cross-term range.
template <class C, int NW>
__device__ __forceinline__ void restage_ct_raw(float* __restrict__ lds,
                                               const float* __restrict__ L, int zdz, int xs,
                                               int ys, int nx, int ny, int nz, long long plane,
                                               int wv, int lane) {
    using X = XtCfg<C>;
    constexpr int RPW = (C::RX + NW - 1) / NW;
    int ycl[C::COLCHUNKS];
    bool yin[C::COLCHUNKS];
    lane_columns<C>(ys, ny, lane, ycl, yin);
#pragma unroll 1
    for (int j = 0; j < C::LPLANES; ++j) {
        const int zz = zdz + j;
        const bool zok = zz >= 0 && zz < nz;
        const float* __restrict__ Lz = L + (long long)clampi(zz, nz) * plane;
        float lv[RPW][C::COLCHUNKS];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rr = wv + r * NW;
            const int xx = xs + rr;
            const bool xok = zok && rr < C::RX && xx >= 0 && xx < nx;
            const float* __restrict__ Lr = Lz + (long long)clampi(xx, nx) * ny;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                const float raw = Lr[ycl[cc]];
                lv[r][cc] = (xok && yin[cc]) ? raw : 0.f;
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rr = wv + r * NW;
#pragma unroll
            for (int cc = 0; cc < C::COLCHUNKS; ++cc) {
                const int c = lane + cc * 64;
                if (rr < C::RX && c < C::CW) lds[X::le(j, rr * C::CW + c)] = lv[r][cc];
            }
        }
    }
}

// One work item in cross-term form where the value range allows it (see
// above); the same partial sums as tiled_item within the FP32 tolerance.
template <int W, int BX, int BZ, int NW, int HW>
__device__ __forceinline__ void tiled_item_xt(const TiledArgs& args, int item,
                                              float* __restrict__ lds) {
    using C = TiledCfg<W, BX, BZ, NW, HW>;
    using X = XtCfg<C>;
    const int g = item / args.n_tiles;
    int t = item - g * args.n_tiles;
    const int ty = t % args.tiles_y;
    t /= args.tiles_y;
    const int tx = t % args.tiles_x;
    const int tz = t / args.tiles_x;

    const int tid = threadIdx.x;
    const int wv = tid >> 6;
    const int lane = tid & 63;
    const int seg = lane / W;
    const int ly = lane % W;
    const int sg = wv * C::S + seg;

    const int nx = args.nx, ny = args.ny, nz = args.nz, r = args.r;
    const long long plane = (long long)nx * ny;
    const int y0 = r + ty * C::OY;
    const int x0 = r + tx * C::WX;
    const int z0 = r + tz * BZ;
    const int y = y0 - 1 + ly;
    const int xb = x0 + sg * BX;

    float tv[BZ + 2][BX + 2];
    load_target_block<BX, BZ>(args.img, args.scale, z0, xb, y, nx, ny, nz, plane, tv);
    int big = 0;
#pragma unroll
    for (int pz = 0; pz < BZ + 2; ++pz)
#pragma unroll
        for (int i = 0; i < BX + 2; ++i) big |= !(fabsf(tv[pz][i]) <= kXtBound);
    const bool squared_item = __syncthreads_or(big) != 0;

    float P[BZ][BX], S[BZ][BX];
#pragma unroll
    for (int j = 0; j < BZ; ++j)
#pragma unroll
        for (int i = 0; i < BX; ++i) P[j][i] = S[j][i] = 0.f;

    const int lane_base = sg * BX * C::CW + ly;
    const int xs = x0 - 1 - HW;
    const int ys = y0 - 1 - HW;
    const int a_begin = g * args.group;
    const int a_end = min(args.n_atlas, a_begin + args.group);
    // 2^-TT per output (1 where the item keeps the squared differences)
    float f[BZ][BX];

    const bool live = wave_has_rows<C, BX>(x0, nx - r, wv);

    if (squared_item) {
        for (int a = a_begin; a < a_end; ++a) {
            const float* __restrict__ A = args.atlas + (long long)a * args.N;
            const float* __restrict__ L = args.label + (long long)a * args.N;
            for (int dz = -HW; dz <= HW; ++dz) {
                __syncthreads();
                stage_slab<C, NW>(lds, A, L, args.scale, z0 + dz, xs, ys, nx, ny, nz, plane, wv,
                                  lane);
                __syncthreads();
                if (!live) continue;
#pragma unroll 1
                for (int dy = 0; dy <= 2 * HW; ++dy)
                    tiled_dx_sweep<C, BX, BZ, HW>(lds, lane_base + dy, tv, P, S);
            }
        }
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) f[j][i] = 1.f;
    } else {
        // TT per output: 3x3 sums of T'^2 in (z, x), then over y across lanes
        float tt[BZ][BX];
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                float s = 0.f;
#pragma unroll
                for (int pz = j; pz < j + 3; ++pz)
#pragma unroll
                    for (int px = i; px < i + 3; ++px) s = fmaf(tv[pz][px], tv[pz][px], s);
                tt[j][i] = lane_next(s) + (lane_prev(s) + s);
            }
#pragma unroll
        for (int pz = 0; pz < BZ + 2; ++pz)
#pragma unroll
            for (int i = 0; i < BX + 2; ++i) tv[pz][i] *= 2.f;
        const unsigned lds0 = lds_byte_address(lds);
        for (int a = a_begin; a < a_end; ++a) {
            const float* __restrict__ A = args.atlas + (long long)a * args.N;
            const float* __restrict__ L = args.label + (long long)a * args.N;
            const f32x2* __restrict__ LE = args.le + (long long)a * args.N;
            for (int dz = -HW; dz <= HW; ++dz) {
                __syncthreads();
                const int flag = stage_slab_xt<C, NW>(lds, A, LE, args.scale, z0 + dz, xs, ys, nx,
                                                      ny, nz, plane, wv, lane);
                if (__syncthreads_or(flag) == 0) {
                    if (!live) continue;
#pragma unroll 1
                    for (int dy = 0; dy <= 2 * HW; ++dy) {
                        unsigned b[X::NBS];
                        make_bases_sweep<C>(lds0, lane_base + dy, b);
                        xterm_dx_sweep<C, BX, BZ, HW>(b, tv, P, S);
                    }
                } else {
                    restage_ct_raw<C, NW>(lds, L, z0 + dz, xs, ys, nx, ny, nz, plane, wv, lane);
                    __syncthreads();
                    if (!live) continue;
#pragma unroll 1
                    for (int dx = 0; dx <= 2 * HW; ++dx) {
#pragma unroll 1
                        for (int dy = 0; dy <= 2 * HW; ++dy) {
                            unsigned b[X::NB];
                            make_bases_xt<C>(lds0, lane_base + dx * C::CW + dy, b);
                            scaled_ssd_step<C, BX, BZ>(b, tv, tt, P, S);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) f[j][i] = __builtin_amdgcn_exp2f(-tt[j][i]);
    }

    if (ly >= 1 && ly <= W - 2 && y < ny - r) {
        float* __restrict__ wsP = args.ws + (long long)(2 * g) * args.N;
        float* __restrict__ wsS = wsP + args.N;
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                const int z = z0 + j, x = xb + i;
                if (z < nz - r && x < nx - r) {
                    const long long idx = z * plane + (long long)x * ny + y;
                    wsP[idx] = P[j][i] * f[j][i];
                    wsS[idx] = S[j][i] * f[j][i];
                }
            }
    }
}

template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) void fuse_tiled_xt(TiledArgs args) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int item = xcd_contiguous(blockIdx.x, args.n_items);
    if (args.only && !args.only[item]) return;   // uniform over the workgroup
    tiled_item_xt<W, BX, BZ, NW, HW>(args, item, lds);
}

// ---------------------------------------------------------------------------
// host helpers shared by pct_fuse.hip and pct_batch.hip (defined in pct_fuse.hip)
// ---------------------------------------------------------------------------

constexpr int kLdsPerCu = 160 * 1024;

// The tiled shapes compiled into the library; index = pct_tile_config.id.
const std::vector<TileShape>& shapes();

// Geometry of `shape_id` for a problem (false if the shape/window has no
// kernel or does not fit the LDS).
bool make_geometry(const pct_problem* pb, int shape_id, int group, TileGeometry* g);

// Cost model: time of one work item of `shape_id` (relative units, the same
// for every shape), or -1 where the shape has no geometry for the problem.
double model_item_time(const pct_problem* pb, int shape_id, int group, TileGeometry* geo);

// Tile shape that measured fastest for a half window, or -1.
int preferred_shape(int hw);

constexpr int kNumCu = 256;

// Checks that the current device is a gfx950; returns its ordinal.
int check_device(int* dev_out);

// The library's per-device workspace, grown to at least `bytes`.
int ensure_workspace(int dev, size_t bytes, void** ptr);

// Bytes of workspace a tiled run of `shape_id` with `n_groups` atlas groups
// needs: the partial sums, then for a cross-term shape the (L * 2^-AA, 2^-AA)
// pair of every atlas voxel (at tiled_le_offset).
size_t tiled_ws_bytes(const pct_problem* pb, int n_groups, int shape_id);
size_t tiled_le_offset(const pct_problem* pb, int n_groups);

// Bytes of workspace the device post chain needs for a volume, and the chain
// itself on a caller's workspace of at least that size (pct_post.hip).
size_t postprocess_ws_bytes(int nh, int nx, int ny, int nz);
int postprocess_dev_ws(int nh, int nx, int ny, int nz, float* d_arr, void* ws, hipStream_t s);

// Whether `shape_id` is a cross-term shape (its kernel reads TiledArgs::le).
bool is_xterm_shape(int shape_id);

// Enqueues atlas_sq_box for atlases [0, n_atlas) of `atlas` and their CT
// `label`: the (L * 2^-AA, 2^-AA) pairs into `le`.
int launch_atlas_pairs(const pct_problem* pb, const float* atlas, const float* label,
                       int n_atlas, float scale, f32x2* le, hipStream_t stream);

TiledArgs make_tiled_args(const pct_problem* pb, const TileGeometry& g, const float* img,
                          const float* atlas, const float* label, float* ws);

// Enqueues the finalize pass (sum of group partials, NaN rule, zero border).
int launch_finalize(const pct_problem* pb, const TileGeometry& g, const float* ws, float* out,
                    hipStream_t stream);

// Launches the single-problem tiled kernel of `shape_id` over the items of
// `a` (honouring a.only), without the finalize pass.
int launch_tiled_items(const TiledArgs& a, int shape_id, int hw, hipStream_t stream);

// Records a pair of HIP events on `s` around one kernel launch (the scope of
// the object) while kernel timing is on (pct_kernel_timing); a no-op
// otherwise.
struct TimedLaunch {
    hipEvent_t begin = nullptr, end = nullptr;
    hipStream_t s;
    explicit TimedLaunch(hipStream_t st);
    ~TimedLaunch();
    TimedLaunch(const TimedLaunch&) = delete;
    TimedLaunch& operator=(const TimedLaunch&) = delete;
};

// Enqueues the direct kernel (any window).
int launch_direct(const pct_problem* pb, const float* img, const float* atlas,
                  const float* label, float* out, hipStream_t stream);

// Destroys the cached batch graphs (they hold workspace addresses).
void release_batch_graphs();

// Frees the per-device copies of the embedding network (pct_learned.hip).
void release_learned_models();

// Frees the device buffers, streams and events of the host-buffer entry
// points (pct_fuse.hip).
void release_host_buffers();

}  // namespace pct

