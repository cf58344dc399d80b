// Learned candidate proposal for gfx950: embedding MLP on MFMA, per-tile
// candidate selection, exact FP32 weighting of the kept candidates with the
// tiled box-sum machinery, a tail term for the rest, and the full search as
// fallback.  No counterpart in the reference program, which always weighs
// all n*NH^3 candidates (src/CPU/NewPatchedBrainPureC.cpp:1399-1432).
//
//  embed_kernel   features -> 30x64 -> ReLU -> 64x64 -> ReLU -> 64x16, BF16
//                 MFMA 16x16x16 with FP32 accumulation.  The chain runs
//                 transposed (H^T = W^T X^T, voxels on the N axis): the
//                 16x16 output tile of one layer is, lane for lane, the B
//                 operand of the next layer, so no data moves between layers;
//                 the 28 weight fragments stay in VGPRs.
//  embed_windows_kernel  the same chain over only the voxels the selection
//                 reads; biases and layer-2 weights in LDS (more waves).
//  box_xyz        NH-wide box sums of the atlas CT in one pass (per plane an
//                 LDS pass over y and x, then running sums over z).
//  select_kernel  per (tile, atlas): embedding distance of every window
//                 offset summed over a 4x2x2 probe cluster at the tile centre
//                 (probe-by-voxel dot products on BF16 MFMA, one window
//                 plane at a time), the topk offsets kept (radix select; a
//                 bitonic sort where a positive margin needs the order
//                 statistics), a low score gap flags the item for the full
//                 search.
//  merge_offsets_kernel  per item, the offsets to evaluate in the pruned
//                 kernel's order (kept, then samples, per dz plane).
//  pruned_kernel  per (tile, atlas): the kept offsets and a 3x3x3 lattice of
//                 sample offsets, weighted exactly with the same LDS slabs and
//                 separable box sums as the plain tiled kernel; the other
//                 candidates contribute (CT window sum - CT of the evaluated
//                 candidates) times the mean exact weight of the samples.
//  fallback       the plain tiled kernel over the flagged items only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pct_internal.h"
#include "pct_mlp.h"
#include "pct_tiled.h"

namespace pct {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kFrags = 28;         // 8 (layer 1) + 16 (layer 2) + 4 (layer 3)
constexpr int kSamples = 27;       // 3x3x3 lattice of tail samples
constexpr int kMaxTopk = 64;
constexpr int kMaxOffsets = kMaxTopk + kSamples;

// First coordinate of a tile's probe cluster along one axis: the cluster of
// c voxels (4 along y, 2 along x and z) sits at the centre of the tile's part
// of the interior [r, n - r), moved inside when the tile is thin.
__host__ __device__ __forceinline__ int probe_origin(int t, int ext, int r, int n, int c) {
    const int p0 = r + t * ext;
    const int pe = min(p0 + ext, n - r);
    return max(r, min((p0 + pe) / 2 - c / 2, n - r - c));
}

// ---------------------------------------------------------------------------
// embedding
// ---------------------------------------------------------------------------

typedef short v2s __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

constexpr int kEmbSeg = 256;              // voxels along y per wave job
constexpr int kEmbRow = kEmbSeg + 2;      // staged row length (one-voxel halo)
constexpr int kEmbRows = 10;              // 3x3 neighbour rows + one zero row
constexpr int kEmbWaves = 4;

struct EmbedArgs {
    const float* vol;               // [nvol][nz][nx][ny]
    unsigned short* emb;            // [nvol][nz][nx][ny][16] BF16
    const v4s* frag;                // [28][64] weight fragments
    const float* bias;              // b1[64], b2[64], b3[16]
    int nx, ny, nz, nvol;
    int nseg;                       // ceil(ny / kEmbSeg)
    int jobs;                       // nvol * nz * nx * nseg row segments
    float mu, inv_sd;
};

// Four FP32 values to BF16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ v4s to_bf16(float a, float b, float c, float d) {
    const v2s lo = __builtin_bit_cast(v2s, __builtin_convertvector((f2){a, b}, bf2));
    const v2s hi = __builtin_bit_cast(v2s, __builtin_convertvector((f2){c, d}, bf2));
    return (v4s){lo[0], lo[1], hi[0], hi[1]};
}

// ReLU then BF16 of four accumulators.  Rounding commutes with the clamp,
// and a BF16 word read as int16 is negative exactly when the value is, so
// the clamp is one packed 16-bit integer max per pair.
__device__ __forceinline__ v4s relu_bf16(v4f h) {
    const v4s r = to_bf16(h[0], h[1], h[2], h[3]);
    const v2s zero = {0, 0};
    const v2s lo = __builtin_elementwise_max((v2s){r[0], r[1]}, zero);
    const v2s hi = __builtin_elementwise_max((v2s){r[2], r[3]}, zero);
    return (v4s){lo[0], lo[1], hi[0], hi[1]};
}

// One wave embeds one row segment (volume, z, x, up to kEmbSeg voxels along
// y).  It first stages the normalised values g of the 3x3 neighbouring rows
// (z-1..z+1, x-1..x+1) with a one-voxel y halo in LDS, zero outside the
// volume, plus a zero row; feature k of the voxel at local column yl is then
// sc_k * (row[oa_k + yl] - row[ob_k + yl]) with per-lane constants.  The
// MFMA chain handles 16 voxels per step: lane l holds voxel n = l % 16 and
// feature / hidden group g = l / 16.
__global__ __launch_bounds__(kEmbWaves * 64) void embed_kernel(EmbedArgs a) {
    __shared__ float rows[kEmbWaves][kEmbRows * kEmbRow];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    v4s w1[4][2], w2[4][4], w3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int s = 0; s < 2; ++s) w1[t][s] = a.frag[(t * 2 + s) * 64 + lane];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) w2[t][s] = a.frag[(8 + t * 4 + s) * 64 + lane];
#pragma unroll
    for (int s = 0; s < 4; ++s) w3[s] = a.frag[(24 + s) * 64 + lane];
    v4f b1[4], b2[4], b3;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b1[t][j] = a.bias[16 * t + 4 * g + j];
            b2[t][j] = a.bias[64 + 16 * t + 4 * g + j];
        }
#pragma unroll
    for (int j = 0; j < 4; ++j) b3[j] = a.bias[128 + 4 * g + j];

    // gather plan of the lane's 8 features k = 16 s + 4 g + j
    constexpr int kZero = 9 * kEmbRow;
    int oa[2][4], ob[2][4];
    float sc[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * s + 4 * g + j;
            int o1 = kZero, o2 = kZero;
            float c = 0.f;
            if (k < 27) {                 // row 3 pz + px, column offset py
                o1 = (k / 3) * kEmbRow + k % 3;
                c = 1.f;
            } else if (k == 27) {         // z: rows (2,1) and (0,1)
                o1 = 7 * kEmbRow + 1;
                o2 = 1 * kEmbRow + 1;
                c = 0.5f;
            } else if (k == 28) {         // x: rows (1,2) and (1,0)
                o1 = 5 * kEmbRow + 1;
                o2 = 3 * kEmbRow + 1;
                c = 0.5f;
            } else if (k == 29) {         // y: row (1,1), columns +1 and -1
                o1 = 4 * kEmbRow + 2;
                o2 = 4 * kEmbRow + 0;
                c = 0.5f;
            }
            oa[s][j] = o1;
            ob[s][j] = o2;
            sc[s][j] = c;
        }

    float* __restrict__ R = rows[wv];
    const long long vsize = (long long)a.nx * a.ny * a.nz;
    for (int base = blockIdx.x * kEmbWaves; base < a.jobs; base += gridDim.x * kEmbWaves) {
        const int job = base + wv;
        const bool active = job < a.jobs;
        int seg = 0, x = 0, z = 0, vi = 0;
        if (active) {
            seg = job % a.nseg;
            int r = job / a.nseg;
            x = r % a.nx;
            r /= a.nx;
            z = r % a.nz;
            vi = r / a.nz;
        }
        const int y0 = seg * kEmbSeg;
        const int len = min(kEmbSeg, a.ny - y0);
        __syncthreads();   // the previous job's reads of R are done
        if (active) {
            const float* __restrict__ v = a.vol + vi * vsize;
#pragma unroll
            for (int rr = 0; rr < kEmbRows; ++rr) {
                const int zz = z + rr / 3 - 1, xx = x + rr % 3 - 1;
                const bool rok = rr < 9 && zz >= 0 && zz < a.nz && xx >= 0 && xx < a.nx;
                const float* __restrict__ src = v + ((long long)zz * a.nx + xx) * a.ny;
                for (int c = lane; c < kEmbRow; c += 64) {
                    const int yy = y0 - 1 + c;
                    float t = 0.f;
                    if (rok && c < len + 2 && yy >= 0 && yy < a.ny) t = (src[yy] - a.mu) * a.inv_sd;
                    R[rr * kEmbRow + c] = t;
                }
            }
        }
        __syncthreads();
        if (!active) continue;
        unsigned short* __restrict__ dst0 =
            a.emb + (vi * vsize + ((long long)z * a.nx + x) * a.ny + y0) * 16 + 4 * g;
        for (int c0 = 0; c0 < len; c0 += 16) {
            const int yl = c0 + n;
            float f[2][4];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    f[s][j] = sc[s][j] * (R[oa[s][j] + yl] - R[ob[s][j] + yl]);
            const v4s xb0 = to_bf16(f[0][0], f[0][1], f[0][2], f[0][3]);
            const v4s xb1 = to_bf16(f[1][0], f[1][1], f[1][2], f[1][3]);
            v4s h1b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4f acc = b1[t];
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[t][0], xb0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[t][1], xb1, acc, 0, 0, 0);
                h1b[t] = relu_bf16(acc);
            }
            v4s h2b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4f acc = b2[t];
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[t][s], h1b[s], acc, 0, 0, 0);
                h2b[t] = relu_bf16(acc);
            }
            v4f e = b3;
#pragma unroll
            for (int s = 0; s < 4; ++s)
                e = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3[s], h2b[s], e, 0, 0, 0);
            if (yl < len) {
                const v4s eb = to_bf16(e[0], e[1], e[2], e[3]);
                *reinterpret_cast<uint2*>(dst0 + (long long)yl * 16) = __builtin_bit_cast(uint2, eb);
            }
        }
    }
}

constexpr int kWinRow = 36;   // LDS row length of the window kernel (>= 4 + 2*7 + 2 + 16)

struct EmbedWinArgs {
    const float* vol;               // [nvol][nz][nx][ny]
    unsigned short* emb;            // [nvol][nz][nx][ny][16] BF16 (window voxels written)
    const v4s* frag;
    const float* bias;
    int nx, ny, nz, nvol;
    int r, hw;
    int tiles_x, tiles_y, oy, wx;   // tile geometry of the selection
    int zlo, nzr;                   // z range [zlo, zlo + nzr) covering every window
    const int* zlist;               // or, when set, only these nzl planes
    int nzl;
    int jobs;                       // nvol * tiles_x * tiles_y * (nzl or nzr)
    float mu, inv_sd;
};

// Embeddings of the atlas voxels the selection reads: for every tile column
// (tx, ty) the probe window's x and y range, over every z any window of the
// column reaches.  One wave per (volume, tile column, z): it stages the 3 x
// (window rows + 2) neighbour rows in LDS and runs the MFMA chain of
// embed_kernel on 16 voxels at a time, with the same feature arithmetic, so
// the values equal embed_kernel's bit for bit.
__global__ __launch_bounds__(kEmbWaves * 64) void embed_windows_kernel(EmbedWinArgs a) {
    __shared__ float rows[kEmbWaves][3 * 18 * kWinRow];
    __shared__ __align__(16) float sbias[144];   // b1[64], b2[64], b3[16]
    __shared__ v4s sw2[16][64];                  // layer-2 weight fragments (t, s) by lane
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    v4s w1[4][2], w3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int s = 0; s < 2; ++s) w1[t][s] = a.frag[(t * 2 + s) * 64 + lane];
#pragma unroll
    for (int s = 0; s < 4; ++s) w3[s] = a.frag[(24 + s) * 64 + lane];
    // the biases stay in LDS and are read per row (128-bit reads), which
    // keeps 36 VGPRs free for more waves per SIMD
    for (int i = threadIdx.x; i < 144; i += kEmbWaves * 64) sbias[i] = a.bias[i];
    for (int i = threadIdx.x; i < 16 * 64; i += kEmbWaves * 64) sw2[i >> 6][i & 63] = a.frag[8 * 64 + i];
    __syncthreads();
    auto bias4 = [&](int i) {   // opaque: read where used, not hoisted out of the loop
        return *reinterpret_cast<const v4f*>(&sbias[opaque(i)]);
    };

    float* __restrict__ R = rows[wv];
    const long long vsize = (long long)a.nx * a.ny * a.nz;
    for (int base = blockIdx.x * kEmbWaves; base < a.jobs; base += gridDim.x * kEmbWaves) {
        const int job = base + wv;
        const bool active = job < a.jobs;
        int z = 0, tx = 0, ty = 0, vi = 0;
        if (active) {
            const int nzj = a.zlist ? a.nzl : a.nzr;
            z = a.zlist ? a.zlist[job % nzj] : a.zlo + job % nzj;
            int t = job / nzj;
            ty = t % a.tiles_y;
            t /= a.tiles_y;
            tx = t % a.tiles_x;
            vi = t / a.tiles_x;
        }
        const int px0 = probe_origin(tx, a.wx, a.r, a.nx, 2);
        const int py0 = probe_origin(ty, a.oy, a.r, a.ny, 4);
        const int xa = max(px0 - a.hw, 0), xb = min(px0 + 1 + a.hw, a.nx - 1);
        const int ya = max(py0 - a.hw, 0), yb = min(py0 + 3 + a.hw, a.ny - 1);
        const int nxr = xb - xa + 3, leny = yb - ya + 1;
        __syncthreads();   // the previous job's reads of R are done
        if (active) {
            const float* __restrict__ v = a.vol + vi * vsize;
            const int cols = leny + 2, total = 3 * nxr * cols;
            for (int e = lane; e < total; e += 64) {
                const int q = e / cols, c = e - (e / cols) * cols;
                const int zz = z + q / nxr - 1, xx = xa - 1 + q % nxr, yy = ya - 1 + c;
                float t = 0.f;
                if (zz >= 0 && zz < a.nz && xx >= 0 && xx < a.nx && yy >= 0 && yy < a.ny)
                    t = (v[((long long)zz * a.nx + xx) * a.ny + yy] - a.mu) * a.inv_sd;
                R[q * kWinRow + c] = t;
            }
        }
        __syncthreads();
        if (!active) continue;
        // gather plan: f = sc * R[oa + o] - sd * R[ob + o], o = xl * kWinRow + yl
        int oa[2][4], ob[2][4];
        float sc[2][4], sd[2][4];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * s + 4 * g + j;
                int o1 = 0, o2 = 0;
                float c1 = 0.f, c2 = 0.f;
                if (k < 27) {
                    o1 = ((k / 9) * nxr + (k / 3) % 3) * kWinRow + k % 3;
                    c1 = 1.f;
                } else if (k == 27) {
                    o1 = (2 * nxr + 1) * kWinRow + 1;
                    o2 = (0 * nxr + 1) * kWinRow + 1;
                    c1 = c2 = 0.5f;
                } else if (k == 28) {
                    o1 = (1 * nxr + 2) * kWinRow + 1;
                    o2 = (1 * nxr + 0) * kWinRow + 1;
                    c1 = c2 = 0.5f;
                } else if (k == 29) {
                    o1 = (1 * nxr + 1) * kWinRow + 2;
                    o2 = (1 * nxr + 1) * kWinRow + 0;
                    c1 = c2 = 0.5f;
                }
                oa[s][j] = o1;
                ob[s][j] = o2;
                sc[s][j] = c1;
                sd[s][j] = c2;
            }
        for (int xl = 0; xl <= xb - xa; ++xl) {
            unsigned short* __restrict__ dst0 =
                a.emb + (vi * vsize + ((long long)z * a.nx + xa + xl) * a.ny + ya) * 16 + 4 * g;
            for (int c0 = 0; c0 < leny; c0 += 16) {
                const int yl = c0 + n;
                const int o = xl * kWinRow + yl;
                float f[2][4];
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        f[s][j] = fmaf(-sd[s][j], R[ob[s][j] + o], sc[s][j] * R[oa[s][j] + o]);
                const v4s xb0 = to_bf16(f[0][0], f[0][1], f[0][2], f[0][3]);
                const v4s xb1 = to_bf16(f[1][0], f[1][1], f[1][2], f[1][3]);
                v4s h1b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    v4f acc = bias4(16 * t + 4 * g);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[t][0], xb0, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[t][1], xb1, acc, 0, 0, 0);
                    h1b[t] = relu_bf16(acc);
                }
                v4s h2b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    v4f acc = bias4(64 + 16 * t + 4 * g);
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            sw2[opaque(t * 4 + s)][lane], h1b[s], acc, 0, 0, 0);
                    h2b[t] = relu_bf16(acc);
                }
                v4f e = bias4(128 + 4 * g);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    e = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3[s], h2b[s], e, 0, 0, 0);
                if (yl < leny) {
                    const v4s eb = to_bf16(e[0], e[1], e[2], e[3]);
                    *reinterpret_cast<uint2*>(dst0 + (long long)yl * 16) =
                        __builtin_bit_cast(uint2, eb);
                }
            }
        }
    }
}

// Weight fragments in MFMA A-operand order: fragment (layer, M-tile t,
// K-step s), lane l, element j holds W[k = 16 s + 4 (l / 16) + j][m = 16 t + l % 16].
static void make_fragments(const float* p, std::vector<unsigned short>& frag,
                           std::vector<float>& bias) {
    auto bf = [](float f) {
        unsigned u;
        std::memcpy(&u, &f, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    frag.assign(kFrags * 64 * 4, 0);
    auto put = [&](int fi, const float* W, int ncols, int t, int s) {
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * s + 4 * (l / 16) + j, m = 16 * t + l % 16;
                frag[(fi * 64 + l) * 4 + j] = bf(W[k * ncols + m]);
            }
    };
    for (int t = 0; t < 4; ++t)
        for (int s = 0; s < 2; ++s) put(t * 2 + s, p + kOffW1, 64, t, s);
    for (int t = 0; t < 4; ++t)
        for (int s = 0; s < 4; ++s) put(8 + t * 4 + s, p + kOffW2, 64, t, s);
    for (int s = 0; s < 4; ++s) put(24 + s, p + kOffW3, 16, 0, s);
    bias.assign(144, 0.f);
    for (int i = 0; i < 64; ++i) bias[i] = p[kOffB1 + i];
    for (int i = 0; i < 64; ++i) bias[64 + i] = p[kOffB2 + i];
    for (int i = 0; i < 16; ++i) bias[128 + i] = p[kOffB3 + i];
}

constexpr size_t kModelBytes = kFrags * 64 * 8 + 144 * 4;

// Per-device copy of the weight fragments and biases, uploaded again only
// when the parameter vector changes.  The upload is ordered on the caller's
// stream; calls that overlap on one device must share a stream (as for the
// workspace).
struct ModelCache {
    std::mutex mu;
    void* d_model[64] = {};
    std::vector<float> last[64];
};
static ModelCache g_model;

// Second stream of the pruned call, per device, created on first use at the
// default priority: the CT window sums depend on neither the embeddings nor
// the selection and run beside them.  The fork
// and join events are recorded and waited on under `mu`, so calls from
// several host threads cannot re-record one between its record and its wait.
struct SideStreams {
    std::mutex mu;
    hipStream_t st[64] = {};
    hipEvent_t fork[64] = {}, join[64] = {};
};
static SideStreams g_side;

static int side_stream(int dev) {   // caller holds g_side.mu
    if (dev < 0 || dev >= 64) return PCT_ERR_NO_DEVICE;
    if (g_side.st[dev]) return PCT_OK;
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (join) (void)hipEventDestroy(join);
        if (fork) (void)hipEventDestroy(fork);
        if (st) (void)hipStreamDestroy(st);
        return PCT_ERR_HIP;
    }
    g_side.st[dev] = st;   // all three, or none
    g_side.fork[dev] = fork;
    g_side.join[dev] = join;
    return PCT_OK;
}

// Per-device copy of the window embedding's z-plane list, uploaded again
// only when the list changes: a copy from pageable memory makes the host
// wait, which would stall every call.
struct ZlistCache {
    std::mutex mu;
    int* d[64] = {};
    size_t cap[64] = {};
    std::vector<int> last[64];
};
static ZlistCache g_zlist;

static int device_zlist(int dev, const std::vector<int>& zl, hipStream_t s, const int** out) {
    if (dev < 0 || dev >= 64) return PCT_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_zlist.mu);
    if (g_zlist.cap[dev] < zl.size()) {
        if (g_zlist.d[dev]) {
            (void)hipStreamSynchronize(s);   // earlier calls may still read it
            (void)hipFree(g_zlist.d[dev]);
        }
        g_zlist.d[dev] = nullptr;
        g_zlist.cap[dev] = 0;
        g_zlist.last[dev].clear();
        if (hipMalloc(&g_zlist.d[dev], zl.size() * sizeof(int)) != hipSuccess) {
            (void)hipGetLastError();
            g_zlist.d[dev] = nullptr;
            return PCT_ERR_NO_MEMORY;
        }
        g_zlist.cap[dev] = zl.size();
    }
    if (g_zlist.last[dev] != zl) {
        // ordered on the caller's stream after the calls that read the old list
        if (hipMemcpyAsync(g_zlist.d[dev], zl.data(), zl.size() * sizeof(int),
                           hipMemcpyHostToDevice, s) != hipSuccess)
            return PCT_ERR_HIP;
        g_zlist.last[dev] = zl;
    }
    *out = g_zlist.d[dev];
    return PCT_OK;
}

void release_learned_models() {
    {
        std::lock_guard<std::mutex> lk(g_zlist.mu);
        for (int dev = 0; dev < 64; ++dev) {
            if (!g_zlist.d[dev]) continue;
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(dev);
            (void)hipDeviceSynchronize();
            (void)hipFree(g_zlist.d[dev]);
            (void)hipSetDevice(cur);
            g_zlist.d[dev] = nullptr;
            g_zlist.cap[dev] = 0;
            g_zlist.last[dev].clear();
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_side.mu);
        for (int dev = 0; dev < 64; ++dev) {
            if (!g_side.st[dev]) continue;
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(dev);
            (void)hipStreamSynchronize(g_side.st[dev]);
            (void)hipStreamDestroy(g_side.st[dev]);
            if (g_side.fork[dev]) (void)hipEventDestroy(g_side.fork[dev]);
            if (g_side.join[dev]) (void)hipEventDestroy(g_side.join[dev]);
            (void)hipSetDevice(cur);
            g_side.st[dev] = nullptr;
            g_side.fork[dev] = g_side.join[dev] = nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(g_model.mu);
    for (int dev = 0; dev < 64; ++dev) {
        if (!g_model.d_model[dev]) continue;
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(dev);
        (void)hipDeviceSynchronize();
        (void)hipFree(g_model.d_model[dev]);
        (void)hipSetDevice(cur);
        g_model.d_model[dev] = nullptr;
        g_model.last[dev].clear();
    }
}

static int device_model(int dev, const float* params, hipStream_t stream, const void** out) {
    if (dev < 0 || dev >= 64) return PCT_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_model.mu);
    if (!g_model.d_model[dev]) {
        if (hipMalloc(&g_model.d_model[dev], kModelBytes) != hipSuccess) {
            (void)hipGetLastError();
            g_model.d_model[dev] = nullptr;
            return PCT_ERR_NO_MEMORY;
        }
        g_model.last[dev].clear();
    }
    std::vector<float>& last = g_model.last[dev];
    if (last.size() != PCT_MLP_NPARAMS ||
        std::memcmp(last.data(), params, sizeof(float) * PCT_MLP_NPARAMS) != 0) {
        std::vector<unsigned short> frag;
        std::vector<float> bias;
        make_fragments(params, frag, bias);
        char* d = static_cast<char*>(g_model.d_model[dev]);
        if (hipMemcpyAsync(d, frag.data(), frag.size() * 2, hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            hipMemcpyAsync(d + frag.size() * 2, bias.data(), bias.size() * 4,
                           hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {   // the host vectors die here
            last.clear();
            return PCT_ERR_HIP;
        }
        last.assign(params, params + PCT_MLP_NPARAMS);
    }
    *out = g_model.d_model[dev];
    return PCT_OK;
}

static int launch_embed(const float* params, const void* d_model, const float* vol, int nx,
                        int ny, int nz, int nvol, unsigned short* emb, hipStream_t stream) {
    EmbedArgs a;
    a.vol = vol;
    a.emb = emb;
    a.frag = static_cast<const v4s*>(d_model);
    a.bias = reinterpret_cast<const float*>(static_cast<const char*>(d_model) + kFrags * 64 * 8);
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.nvol = nvol;
    a.nseg = (ny + kEmbSeg - 1) / kEmbSeg;
    const long long jobs = (long long)nvol * nz * nx * a.nseg;
    if (jobs >= INT_MAX) return PCT_ERR_TOO_LARGE;
    a.jobs = (int)jobs;
    a.mu = params[kOffMu];
    a.inv_sd = params[kOffSd];
    const long long blocks = std::min<long long>((jobs + kEmbWaves - 1) / kEmbWaves, 256LL * 16);
    if (blocks > 0)
        hipLaunchKernelGGL(embed_kernel, dim3((unsigned)blocks), dim3(kEmbWaves * 64), 0, stream, a);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

// Embeddings inside the selection's probe windows (see embed_windows_kernel):
// `nvol` volumes, window half-width `whw` around each probe cluster (the
// search half window for the atlases, 0 for the target, whose probe
// clusters alone are read); tile geometry as in pct_fuse_pruned_dev.
static int launch_embed_windows(const float* params, const void* d_model, const float* vol,
                                int nvol, int whw, const pct_problem* pb, const TileGeometry& g,
                                int oy, int wx, int bz, unsigned short* emb, hipStream_t stream,
                                const int* d_zlist = nullptr, int nzl = 0) {
    EmbedWinArgs a;
    a.vol = vol;
    a.emb = emb;
    a.frag = static_cast<const v4s*>(d_model);
    a.bias = reinterpret_cast<const float*>(static_cast<const char*>(d_model) + kFrags * 64 * 8);
    a.nx = pb->nx;
    a.ny = pb->ny;
    a.nz = pb->nz;
    a.nvol = nvol;
    a.r = g.r;
    a.hw = whw;
    a.tiles_x = g.tiles_x;
    a.tiles_y = g.tiles_y;
    a.oy = oy;
    a.wx = wx;
    if (g.n_tiles == 0) return PCT_OK;
    const int z_first = probe_origin(0, bz, g.r, pb->nz, 2);
    const int z_last = probe_origin(g.tiles_z - 1, bz, g.r, pb->nz, 2);
    a.zlo = std::max(z_first - whw, 0);
    const int zhi = std::min(z_last + 1 + whw, pb->nz - 1);
    a.nzr = zhi - a.zlo + 1;
    a.zlist = d_zlist;
    a.nzl = nzl;
    const long long jobs = (long long)a.nvol * a.tiles_x * a.tiles_y * (d_zlist ? nzl : a.nzr);
    if (jobs >= INT_MAX) return PCT_ERR_TOO_LARGE;
    a.jobs = (int)jobs;
    a.mu = params[kOffMu];
    a.inv_sd = params[kOffSd];
    const long long blocks = std::min<long long>((jobs + kEmbWaves - 1) / kEmbWaves, 256LL * 16);
    if (blocks > 0)
        hipLaunchKernelGGL(embed_windows_kernel, dim3((unsigned)blocks), dim3(kEmbWaves * 64), 0,
                           stream, a);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

// ---------------------------------------------------------------------------
// window box sums of the atlas CT
// ---------------------------------------------------------------------------

// The NH^3 window sums in one pass.  A block owns a kBoxTX x kBoxTY column
// tile of one volume over a run of kBoxRun output planes and sweeps the
// planes that feed it.  Each plane's tile, with a halo of HW (zeros outside
// the volume), goes through LDS and is summed over y (4 outputs per task,
// from 128-bit LDS reads), then over x.  The plane sums enter a running sum
// over z in double, and the last 2 HW + 2 of them stay in registers for the
// subtraction; the double sum of float plane sums is exact in practice, so
// the result does not depend on where a run starts.  Every thread owns 2
// consecutive x of one y of the tile, and the loads of the next two planes
// are in flight during a plane's sums (two register buffers).
constexpr int kBoxTX = 16, kBoxTY = 64, kBoxRun = 32, kBoxThreads = 512;

template <int HW>
__global__ __launch_bounds__(kBoxThreads) void box_xyz(const float* __restrict__ in,
                                                       float* __restrict__ out, int nx, int ny,
                                                       int nz, int nvol) {
    constexpr int K = 2 * HW + 1;
    constexpr int KR = K + 1;                    // ring length, even
    constexpr int R = kBoxTX + 2 * HW, C = kBoxTY + 2 * HW;
    constexpr int NQ = (4 + 2 * HW + 3) / 4;     // float4 reads per y-sum task
    constexpr int CP = kBoxTY - 4 + 4 * NQ;      // LDS row pitch (>= C, multiple of 4)
    constexpr int NLD = (R * C + kBoxThreads - 1) / kBoxThreads;   // tile loads per thread
    constexpr int NYT = R * (kBoxTY / 4);        // y-sum tasks
    constexpr int OX = kBoxTX / (kBoxThreads / 64);   // x outputs per thread
    // row pitch of the y sums: a multiple of 4 for the 128-bit stores, and
    // not of 64, so the column reads of the x sums (two rows per
    // instruction) hit different banks
    constexpr int YP = kBoxTY + 4;
    static_assert(CP >= C && CP % 4 == 0, "row pitch");
    static_assert(NYT <= kBoxThreads && OX == 2, "one y-sum task and 2 outputs per thread");
    __shared__ __align__(16) float T[R * CP];
    __shared__ __align__(16) float Y[R * YP];

    const int tiles_x = (nx + kBoxTX - 1) / kBoxTX, tiles_y = (ny + kBoxTY - 1) / kBoxTY;
    const int runs = (nz + kBoxRun - 1) / kBoxRun;
    long long b = blockIdx.x;
    const int ty = (int)(b % tiles_y);
    b /= tiles_y;
    const int tx = (int)(b % tiles_x);
    b /= tiles_x;
    const int run = (int)(b % runs);
    const long long vi = b / runs;
    if (vi >= nvol) return;   // uniform over the block
    const int x0 = tx * kBoxTX, y0 = ty * kBoxTY;
    const int z0 = run * kBoxRun, z1 = min(nz, z0 + kBoxRun);
    const long long plane = (long long)nx * ny;
    const float* __restrict__ src = in + vi * plane * nz;
    float* __restrict__ dst = out + vi * plane * nz;

    // load slots: in-plane offset of the tile element (0 where it lies
    // outside the volume or past the tile, so that every load is
    // unconditional and the loads of two planes stay in flight: a load under
    // a branch makes the compiler wait for every earlier load), whether it
    // lies in the volume (else a zero is stored), and whether it lies in the
    // tile (else nothing is stored)
    int lofs[NLD];
    bool lin[NLD], lst[NLD];
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int e = (int)threadIdx.x + i * kBoxThreads;
        const int rr = e / C, cc = e - (e / C) * C;
        const int x = x0 - HW + rr, y = y0 - HW + cc;
        lst[i] = e < R * C;
        lin[i] = lst[i] && x >= 0 && x < nx && y >= 0 && y < ny;
        lofs[i] = lin[i] ? x * ny + y : 0;
    }
    float va[NLD], vb[NLD];
    auto load = [&](int p, float (&v)[NLD]) {
        const float* __restrict__ sp = src + p * plane;
#pragma unroll
        for (int i = 0; i < NLD; ++i) v[i] = sp[lofs[i]];
    };

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ox = OX * wv, oy = lane;   // outputs (x0 + ox + m, y0 + oy), m < OX
    float ring[KR][OX];
    double acc[OX] = {0.0, 0.0};
    const int pa = max(0, z0 - HW), pe = z1 - 1 + HW;   // planes >= nz are zero
    load(pa, va);
    load(min(pa + 1, nz - 1), vb);
    for (int pb = pa; pb <= pe; pb += KR) {
#pragma unroll
        for (int j = 0; j < KR; ++j) {
            const int p = pb + j;
            if (p > pe) break;   // uniform
            float(&cur)[NLD] = (j & 1) ? vb : va;   // holds plane p
            float bs[OX] = {0.f, 0.f};
            if (p < nz) {        // uniform
                __syncthreads();   // the previous plane's y sums are done with T
#pragma unroll
                for (int i = 0; i < NLD; ++i) asm volatile("" : "+v"(cur[i]));
#pragma unroll
                for (int i = 0; i < NLD; ++i) {
                    const int e = (int)threadIdx.x + i * kBoxThreads;
                    if (lst[i]) T[(e / C) * CP + e - (e / C) * C] = lin[i] ? cur[i] : 0.f;
                }
            }
            // plane p + 2, in flight for two planes; loaded on every path (a
            // clamped plane where it is not needed), so that the compiler can
            // count the loads in flight and wait for the older plane only
            load(min(p + 2, nz - 1), cur);
            if (p < nz) {        // uniform
                __syncthreads();
                if ((int)threadIdx.x < NYT) {
                    const int rr = threadIdx.x >> 4, c0 = 4 * (threadIdx.x & 15);
                    float w4[4 * NQ];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const float4 f = *reinterpret_cast<const float4*>(&T[rr * CP + c0 + 4 * q]);
                        w4[4 * q] = f.x;
                        w4[4 * q + 1] = f.y;
                        w4[4 * q + 2] = f.z;
                        w4[4 * q + 3] = f.w;
                    }
                    float4 o;
                    float* op = reinterpret_cast<float*>(&o);
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        float sacc = 0.f;
#pragma unroll
                        for (int k = 0; k < K; ++k) sacc += w4[m + k];
                        op[m] = sacc;
                    }
                    *reinterpret_cast<float4*>(&Y[rr * YP + c0]) = o;
                }
                __syncthreads();
                float col[OX + 2 * HW];
#pragma unroll
                for (int k = 0; k < OX + 2 * HW; ++k) col[k] = Y[(ox + k) * YP + oy];
#pragma unroll
                for (int m = 0; m < OX; ++m) {
                    float sacc = 0.f;
#pragma unroll
                    for (int k = 0; k < K; ++k) sacc += col[m + k];
                    bs[m] = sacc;
                }
            }
#pragma unroll
            for (int m = 0; m < OX; ++m) {
                acc[m] += (double)bs[m];
                ring[j][m] = bs[m];
            }
            const int z = p - HW;   // the output whose window ends at p
            if (z >= z0) {
                const int y = y0 + oy;
#pragma unroll
                for (int m = 0; m < OX; ++m) {
                    const int x = x0 + ox + m;
                    if (x < nx && y < ny) dst[z * plane + (long long)x * ny + y] = (float)acc[m];
                    // plane z - HW = p - 2 HW leaves the window; it sits in
                    // ring slot (j - 2 HW) mod KR = (j + 2) mod KR
                    if (z - HW >= pa) acc[m] -= (double)ring[(j + 2) % KR][m];
                }
            }
        }
    }
}

using BoxFn = void (*)(const float*, float*, int, int, int, int);
static const BoxFn kBoxXyz[kMaxTiledHalfWindow] = {&box_xyz<1>, &box_xyz<2>, &box_xyz<3>,
                                                   &box_xyz<4>, &box_xyz<5>, &box_xyz<6>,
                                                   &box_xyz<7>};

// ---------------------------------------------------------------------------
// candidate selection
// ---------------------------------------------------------------------------

struct SelectArgs {
    const unsigned short* embT;   // [N][16]
    const unsigned short* embA;   // [n][N][16]
    int* sel;                     // [items][topk] kept offset indices, ascending, kNoOffset after
    unsigned char* flag;          // [items] 1 = full search
    long long N;
    int nx, ny, nz, hw, r;
    int tiles_y, tiles_x, n_tiles;
    int oy, wx, bz;               // tile extents in y, x, z
    int topk;
    float margin;
    int max_slabs;                // distinct dz slabs the kept offsets may use; 0 = all
    int n_lattice_dz;
    int lattice_dz[3];            // dz of the lattice samples (always allowed)
    unsigned score_bits;          // dz indices (dz + hw) whose offsets are scored
    unsigned fixed_allowed;       // the allowed dz indices when they do not depend on the
                                  // scores, else 0
};

constexpr int kSelThreads = 256;
constexpr int kSelPerThread = 16;   // 4096 offsets (NH = 15: 3375)
static_assert(kSelThreads >= 16 * 15, "16 threads per dz slab in the slab minima");
static_assert(kSelThreads >= 15 * 15, "one thread per (dx, dy) in the gathers");
constexpr int kNoOffset = 1 << 30;

// One workgroup per item = (atlas, tile), item = atlas * n_tiles + tile.
// Order-preserving integer image of a float score (-0 is taken as +0, as the
// float comparison does).
__device__ __forceinline__ unsigned order_key(float f) {
    const unsigned u = __float_as_uint(f + 0.f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// The topk smallest (score, offset index) pairs among the m candidates of the
// allowed slabs, written in ascending offset order, and the item's flag for a
// margin <= 0 (set for a negative margin).  Thread t holds the scores of
// offsets t + q kSelThreads.  A four-pass radix select over the order keys
// finds T, the key of the k-th smallest; keys below T are kept, and of the
// keys equal to T the ones with the smallest offsets.  Ordered compaction by
// wave ballots writes the list.  The set equals the first k entries of the
// full sort by (score, offset index).
__device__ __forceinline__ void select_topk_radix(
    const SelectArgs& a, int item, int ncand, int nw2, const float (&sc)[kSelPerThread],
    const float* key, unsigned allowed_bits, int m, unsigned* hist, unsigned& sh_T,
    int& sh_need, int& sh_cnt, int* tie, int* wcnt) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = min(a.topk, m);
    bool al[kSelPerThread];
    unsigned uk[kSelPerThread];
#pragma unroll
    for (int q = 0; q < kSelPerThread; ++q) {
        const int c = threadIdx.x + q * kSelThreads;
        al[q] = c < ncand && ((allowed_bits >> (c / nw2)) & 1u);
        uk[q] = order_key(sc[q]);
    }
    unsigned T = 0;
    int need = k, n_eq = k;
    if (k < m) {
        for (int shift = 24; shift >= 0; shift -= 8) {
            if (threadIdx.x < 256) hist[threadIdx.x] = 0;
            __syncthreads();
            const unsigned hi_mask = shift == 24 ? 0u : (0xffffffffu << (shift + 8));
#pragma unroll
            for (int q = 0; q < kSelPerThread; ++q)
                if (al[q] && ((uk[q] ^ T) & hi_mask) == 0)
                    atomicAdd(&hist[(uk[q] >> shift) & 255u], 1u);
            __syncthreads();
            if (wv == 0) {
                // lane l holds bins 4l .. 4l+3; the bin where the running
                // count reaches `need` extends T by one byte
                const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1],
                               h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
                const unsigned own = h0 + h1 + h2 + h3;
                unsigned incl = own;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const unsigned v = __shfl_up(incl, o);
                    if (lane >= o) incl += v;
                }
                const unsigned excl = incl - own;
                const unsigned nd = (unsigned)need;
                if (excl < nd && nd <= incl) {
                    unsigned before = excl, cnt = h0;
                    int j = 0;
                    if (before + h0 < nd) {
                        before += h0;
                        j = 1;
                        cnt = h1;
                        if (before + h1 < nd) {
                            before += h1;
                            j = 2;
                            cnt = h2;
                            if (before + h2 < nd) {
                                before += h2;
                                j = 3;
                                cnt = h3;
                            }
                        }
                    }
                    sh_T = T | ((unsigned)(4 * lane + j) << shift);
                    sh_need = (int)(nd - before);
                    sh_cnt = (int)cnt;
                }
            }
            __syncthreads();
            T = sh_T;
            need = sh_need;
            n_eq = sh_cnt;
        }
        if (need < n_eq) {
            // more keys equal T than places: the smallest offsets win
            if (threadIdx.x == 0) {
                int taken = 0;
                for (int c = 0; c < ncand && taken < need; ++c)
                    if (((allowed_bits >> (c / nw2)) & 1u) && order_key(key[c]) == T)
                        tie[taken++] = c;
            }
            __syncthreads();
        }
    }
    bool keep[kSelPerThread];
#pragma unroll
    for (int q = 0; q < kSelPerThread; ++q) {
        const int c = threadIdx.x + q * kSelThreads;
        bool kp = al[q] && (k == m || uk[q] < T);
        if (al[q] && k < m && uk[q] == T) {
            if (need == n_eq) {
                kp = true;
            } else {
                for (int j = 0; j < need; ++j) kp |= tie[j] == c;
            }
        }
        keep[q] = kp;
    }
    int* out = a.sel + (long long)item * a.topk;
    int base = 0;
#pragma unroll
    for (int q = 0; q < kSelPerThread; ++q) {
        if (q * kSelThreads >= ncand) break;   // uniform over the workgroup
        {
            // a round without allowed slabs keeps nothing: skip its barriers
            const int lo = q * kSelThreads / nw2;
            const int hi = (min(ncand, (q + 1) * kSelThreads) - 1) / nw2;
            if (((allowed_bits >> lo) & ((2u << (hi - lo)) - 1u)) == 0u) continue;
        }
        const unsigned long long bal = __ballot(keep[q]);
        if (lane == 0) wcnt[wv] = __popcll(bal);
        __syncthreads();
        int off = base, tot = 0;
#pragma unroll
        for (int w = 0; w < kSelThreads / 64; ++w) {
            if (w < wv) off += wcnt[w];
            tot += wcnt[w];
        }
        if (keep[q])
            out[off + __popcll(bal & ((1ull << lane) - 1ull))] = threadIdx.x + q * kSelThreads;
        __syncthreads();   // wcnt is rewritten for the next q
        base += tot;
    }
    for (int i = k + threadIdx.x; i < a.topk; i += kSelThreads) out[i] = kNoOffset;
    if (threadIdx.x == 0) a.flag[item] = (unsigned char)(a.margin < 0.f ? 1 : 0);
}

__global__ __launch_bounds__(kSelThreads) void select_kernel(SelectArgs a) {
    extern __shared__ float sh[];
    __shared__ float dzmin[16];
    __shared__ int allowed[16];       // allowed dz indices, ascending
    __shared__ int n_allowed;
    __shared__ unsigned allowed_bits;
    __shared__ unsigned rs_hist[256];      // radix-select histogram
    __shared__ unsigned rs_T;
    __shared__ int rs_need, rs_cnt;
    __shared__ int rs_tie[kMaxTopk];
    __shared__ int rs_wcnt[kSelThreads / 64];
    const int item = blockIdx.x;
    const int at = item / a.n_tiles;
    int t = item - at * a.n_tiles;
    const int ty = t % a.tiles_y;
    t /= a.tiles_y;
    const int tx = t % a.tiles_x;
    const int tz = t / a.tiles_x;
    const int r = a.r, hw = a.hw, nw = 2 * hw + 1, nw2 = nw * nw;
    const int ncand = nw2 * nw;
    // probe cluster: 4 (y) x 2 (x) x 2 (z) voxels at the centre of the
    // tile's part of the interior; probe p sits at (p & 3, (p >> 2) & 1,
    // p >> 3).  Its window spans WY x WXX x WZ atlas voxels.  Coordinates are
    // clamped into the volume (a thin interior can leave the window partly
    // outside).
    const int py0 = probe_origin(ty, a.oy, r, a.ny, 4);
    const int px0 = probe_origin(tx, a.wx, r, a.nx, 2);
    const int pz0 = probe_origin(tz, a.bz, r, a.nz, 2);
    const int WY = 4 + 2 * hw, WXX = 2 + 2 * hw, WZ = 2 + 2 * hw;
    const int pv = WY * WXX;                   // voxels per window plane
    const int ntile = (pv + 15) >> 4;          // 16-voxel MFMA tiles per plane
    const int pvs = ntile << 4;                // plane stride of the G ring
    const long long plane = (long long)a.nx * a.ny;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ln = lane & 15, lg = lane >> 4;
    // The score of offset d is sum_p |t_p - a(p + d)|^2 less the common
    // sum_p |t_p|^2, i.e. sum_p G[p][p + d] with G[p][w] = -2 t_p . a_w +
    // |a_w|^2.  Per window plane, one BF16 MFMA 16x16x16 gives -2 t_p . a_w
    // for the 16 probes and 16 voxels; |a_w|^2 is added in FP32.  G of two
    // consecutive planes lives in an LDS ring; the offsets of one dz read the
    // plane of the z = 0 probes and the next one.
    float* key = sh;                                   // [ncand] scores by offset
    float* ring = sh + ((ncand + 63) & ~63);           // [16][pvs] G of one window plane
    v4s aop;   // A operand: lane holds -2 t[p = ln][k = 4 lg .. 4 lg + 3]
    {
        const int yy = min(py0 + (ln & 3), a.ny - 1), xx = min(px0 + ((ln >> 2) & 1), a.nx - 1),
                  zz = min(pz0 + (ln >> 3), a.nz - 1);
        const uint2 raw = *reinterpret_cast<const uint2*>(
            a.embT + (zz * plane + (long long)xx * a.ny + yy) * 16 + 4 * lg);
        const v4s t = __builtin_bit_cast(v4s, raw);
        // -2 t is exact in BF16
        aop = (v4s){(short)f2bf(-2.f * bf2f((unsigned short)t[0])),
                    (short)f2bf(-2.f * bf2f((unsigned short)t[1])),
                    (short)f2bf(-2.f * bf2f((unsigned short)t[2])),
                    (short)f2bf(-2.f * bf2f((unsigned short)t[3]))};
    }
    const unsigned short* __restrict__ eA = a.embA + (long long)at * a.N * 16;
    // B operands of one plane: wave wv takes tiles wv, wv + 8, ...; lane
    // holds dims 4 lg .. 4 lg + 3 of voxel 16 tile + ln
    constexpr int kTilesPerWave = 5;   // 18 tiles at NH = 15 over 4 waves
    // in-plane element offset of the lane's window voxel per tile (the same
    // on every plane), -1 past the window
    long long xyo[kTilesPerWave];
#pragma unroll
    for (int i = 0; i < kTilesPerWave; ++i) {
        const int tile = wv + i * (kSelThreads / 64);
        const int wl = tile * 16 + ln;
        xyo[i] = -1;
        if (tile < ntile && wl < pv) {
            const int wy = wl % WY, wx = wl / WY;
            const int yy = min(max(py0 - hw + wy, 0), a.ny - 1);
            const int xx = min(max(px0 - hw + wx, 0), a.nx - 1);
            xyo[i] = ((long long)xx * a.ny + yy) * 16 + 4 * lg;
        }
    }
    auto load_plane = [&](int wz, uint2 (&b)[kTilesPerWave]) {
        const long long zo = (long long)min(max(pz0 - hw + wz, 0), a.nz - 1) * plane * 16;
#pragma unroll
        for (int i = 0; i < kTilesPerWave; ++i)
            b[i] = xyo[i] >= 0 ? *reinterpret_cast<const uint2*>(eA + zo + xyo[i])
                               : make_uint2(0u, 0u);
    };
    auto plane_g = [&](int slot, const uint2 (&b)[kTilesPerWave]) {
        float* g = ring + slot * 16 * pvs;
#pragma unroll
        for (int i = 0; i < kTilesPerWave; ++i) {
            const int tile = wv + i * (kSelThreads / 64);
            if (tile >= ntile) break;   // uniform over the wave
            const v4s bop = __builtin_bit_cast(v4s, b[i]);
            const v4f d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                aop, bop, (v4f){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            float nn = 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = bf2f((unsigned short)bop[j]);
                nn = fmaf(v, v, nn);
            }
            nn += __shfl_xor(nn, 16);
            nn += __shfl_xor(nn, 32);
            const int wl = tile * 16 + ln;
#pragma unroll
            for (int j = 0; j < 4; ++j) g[(4 * lg + j) * pvs + wl] = d[j] + nn;
        }
    };
    // Scores are computed for the dz planes in a.score_bits (every plane, or
    // only those kept offsets may come from).  Window plane wz holds the
    // voxels that the probes with z = 0 read for dz index wz and the probes
    // with z = 1 read for dz index wz - 1, so a plane is computed when one of
    // those is scored, and one pass over its G finishes the score of dz
    // index wz - 1 and starts that of wz.  A gathering thread keeps its
    // offset's four partial sums in registers from one plane to the next.
    // The loads of the next such plane are in flight during the gathers.
    auto scored = [&](int dzi) { return dzi >= 0 && dzi < nw && ((a.score_bits >> dzi) & 1u); };
    auto plane_needed = [&](int wz) { return scored(wz) || scored(wz - 1); };
    uint2 bq[kTilesPerWave];
    int loaded = -1;   // window plane whose B operands sit in bq (uniform)
    auto ensure = [&](int wz) {
        if (loaded != wz) {
            load_plane(wz, bq);
            loaded = wz;
        }
    };
    const bool gatherer = (int)threadIdx.x < nw2;
    const int dxi = gatherer ? (int)threadIdx.x / nw : 0;
    const int dyi = gatherer ? (int)threadIdx.x - dxi * nw : 0;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int wz = 0; wz < WZ; ++wz) {
        if (!plane_needed(wz)) continue;   // uniform over the workgroup
        ensure(wz);
        plane_g(0, bq);
        for (int w2 = wz + 1; w2 < WZ; ++w2)
            if (plane_needed(w2)) {
                ensure(w2);   // in flight during the gathers
                break;
            }
        __syncthreads();
        if (gatherer) {
            if (scored(wz - 1)) {
                // probes with z = 1 of dz index wz - 1
#pragma unroll
                for (int p = 8; p < 16; ++p)
                    acc[p & 3] += ring[p * pvs + (((p >> 2) & 1) + dxi) * WY + (p & 3) + dyi];
                key[(wz - 1) * nw2 + threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            }
            if (scored(wz)) {
                // probes with z = 0 of dz index wz
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] = 0.f;
#pragma unroll
                for (int p = 0; p < 8; ++p)
                    acc[p & 3] += ring[p * pvs + (((p >> 2) & 1) + dxi) * WY + (p & 3) + dyi];
            }
        }
        __syncthreads();   // the ring is rewritten next
    }
    float sc[kSelPerThread];
#pragma unroll
    for (int q = 0; q < kSelPerThread; ++q) {
        const int c = threadIdx.x + q * kSelThreads;
        sc[q] = c < ncand ? key[c] : __builtin_huge_valf();
    }
    // the allowed slabs: the lattice's and the best others up to max_slabs
    // (ties to the smaller dz).  Where the host has fixed them, every
    // thread below nw places its own slab.
    if (a.fixed_allowed) {   // uniform
        const unsigned bits = a.fixed_allowed;
        if ((int)threadIdx.x < nw && ((bits >> threadIdx.x) & 1u))
            allowed[__popc(bits & ((1u << threadIdx.x) - 1u))] = threadIdx.x;
        if (threadIdx.x == 0) {
            n_allowed = __popc(bits);
            allowed_bits = bits;
        }
        __syncthreads();
    } else {
        // best score of every dz slab (16 threads per slab, then a shuffle
        // reduction)
        {
            const int d = threadIdx.x >> 4, part = threadIdx.x & 15;
            float m = __builtin_huge_valf();
            if (d < nw && scored(d))
                for (int c = d * nw2 + part; c < (d + 1) * nw2; c += 16) m = fminf(m, key[c]);
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off, 16));
            if (d < nw && part == 0) dzmin[d] = m;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            bool ok[16];
            for (int d = 0; d < nw; ++d) ok[d] = a.max_slabs <= 0 || a.max_slabs >= nw;
            int cnt = 0;
            for (int d = 0; d < nw; ++d) cnt += ok[d];
            for (int l = 0; l < a.n_lattice_dz && cnt < nw; ++l) {
                const int d = a.lattice_dz[l] + hw;
                if (!ok[d]) {
                    ok[d] = true;
                    ++cnt;
                }
            }
            while (cnt < a.max_slabs && cnt < nw) {
                int best = -1;
                for (int d = 0; d < nw; ++d)
                    if (!ok[d] && (best < 0 || dzmin[d] < dzmin[best])) best = d;
                ok[best] = true;
                ++cnt;
            }
            int m = 0;
            unsigned bits = 0;
            for (int d = 0; d < nw; ++d)
                if (ok[d]) {
                    allowed[m++] = d;
                    bits |= 1u << d;
                }
            n_allowed = m;
            allowed_bits = bits;
        }
        __syncthreads();
    }
    if (a.margin <= 0.f) {
        // The flag does not depend on the order statistics here, so the kept
        // set - the topk smallest (score, offset index) pairs of the allowed
        // slabs - comes from a radix select, without a sort.
        select_topk_radix(a, item, ncand, nw2, sc, key, allowed_bits, n_allowed * nw2, rs_hist,
                          rs_T, rs_need, rs_cnt, rs_tie, rs_wcnt);
        return;
    }
    // candidates of the allowed slabs, then bitonic sort ascending by
    // (score, offset index)
    const int m = n_allowed * nw2;
    int p2 = 1;
    while (p2 < m) p2 <<= 1;
    float* key2 = sh + ((ncand + 63) & ~63);
    int* val2 = reinterpret_cast<int*>(key2 + p2);
    for (int i = threadIdx.x; i < p2; i += kSelThreads) {
        if (i < m) {
            const int c = allowed[i / nw2] * nw2 + i % nw2;
            key2[i] = key[c];
            val2[i] = c;
        } else {
            key2[i] = __builtin_huge_valf();
            val2[i] = kNoOffset + i;
        }
    }
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < p2 / 2; i += kSelThreads) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const float kl = key2[lo], kh = key2[hi];
                const int vl = val2[lo], vh = val2[hi];
                const bool gt = kl > kh || (kl == kh && vl > vh);
                if (gt == up) {
                    key2[lo] = kh;
                    key2[hi] = kl;
                    val2[lo] = vh;
                    val2[hi] = vl;
                }
            }
            __syncthreads();
        }
    }
    const int k = min(a.topk, m);
    if (threadIdx.x == 0) {
        // a small gap between the last kept and the first dropped score,
        // relative to the spread of the scores, flags the item
        int f = a.margin < 0.f ? 1 : 0;
        if (!f && k < m) {
            const float spread = key2[m / 2] - key2[0];
            const float gap = key2[k] - key2[k - 1];
            f = (spread > 0.f && gap < a.margin * spread) ? 1 : 0;
        }
        a.flag[item] = (unsigned char)f;
    }
    // kept offsets in ascending index order (dz-major), by rank counting
    int* out = a.sel + (long long)item * a.topk;
    for (int i = threadIdx.x; i < a.topk; i += kSelThreads) {
        if (i < k) {
            const int v = val2[i];
            int rank = 0;
            for (int j = 0; j < k; ++j) rank += val2[j] < v;
            out[rank] = v;
        } else {
            out[i] = kNoOffset;
        }
    }
}

// ---------------------------------------------------------------------------
// pruned exact weighting
// ---------------------------------------------------------------------------

struct PrunedArgs {
    TiledArgs t;                    // problem, tiles (group = 1), workspace partials
    const int* sel;                 // [items][topk]
    unsigned char* flag;            // [items]: 1 flagged by the selection, 2 by the error check
    const float* boxL;              // [n][N] window sums of the CT
    const int* merged;              // [items][kMergedStride] offset lists (merge_offsets_kernel)
    int ncand;                      // NH^3
    float max_err;                  // > 0: error bound (label units) above which an item is flagged
};

// Per (tile, atlas) item, the offsets the pruned kernel evaluates, in its
// order: per dz plane (ascending), the kept offsets, then the lattice samples
// that are not kept (a sample that is also kept counts once, as kept).  One
// wave per item, every entry placed by counting: a kept offset goes after the
// kept offsets before it and the samples of the earlier planes; a sample goes
// after the kept offsets of its plane and the planes before, and after the
// samples before it.  Entry 0 holds the count and the number of samples;
// entry 1 + k holds offset k as (sample << 31) | ((dz + 8) << 16) | LDS delta
// (dy + dx rows of the slab, row length cw).
constexpr int kMergeWaves = 4;
constexpr int kMergedStride = 96;   // >= 1 + kMaxOffsets
static_assert(kMergedStride >= 1 + kMaxOffsets, "merged list stride");
static_assert(kMaxTopk <= 64 && kSamples <= 64, "one lane per kept offset and per sample");

struct MergeArgs {
    const int* sel;                 // [items][topk] kept offsets, ascending, kNoOffset after
    int* merged;                    // [items][kMergedStride]
    long long items;
    int topk, hw, cw;
    int n_samples;
    int samples[kSamples];          // lattice offset indices, ascending
};

__global__ __launch_bounds__(kMergeWaves * 64) void merge_offsets_kernel(MergeArgs a) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long item = (long long)blockIdx.x * kMergeWaves + wv;
    if (item >= a.items) return;   // per wave; no workgroup barrier below
    const int nw1 = 2 * a.hw + 1, nw2 = nw1 * nw1;
    const int kv = lane < a.topk ? a.sel[item * a.topk + lane] : kNoOffset;
    int sv = kNoOffset;   // the lane's sample (uniform indices into the arguments)
    for (int j = 0; j < kSamples; ++j)
        if (j == lane && j < a.n_samples) sv = a.samples[j];
    const bool has_k = kv < kNoOffset, has_s = sv < kNoOffset;
    bool dup = false;
    for (int q = 0; q < a.topk; ++q) dup |= __shfl(kv, q) == sv;
    const unsigned long long smask = __ballot(has_s && !dup);   // samples evaluated
    const unsigned long long kmask = __ballot(has_k);
    const int klo = has_k ? (kv / nw2) * nw2 : 0;          // first offset of the kept plane
    const int shi = has_s ? (sv / nw2 + 1) * nw2 : 0;      // end of the sample's plane
    int before_k = 0, before_s = 0;
    for (int j = 0; j < a.n_samples; ++j)
        before_k += (int)((smask >> j) & 1ull) & (__shfl(sv, j) < klo);
    for (int q = 0; q < a.topk; ++q) before_s += __shfl(kv, q) < shi;
    auto pack = [&](int o, unsigned smpl) {
        const int dz = o / nw2 - a.hw;
        const int del = ((o / nw1) % nw1) * a.cw + o % nw1;
        return (int)((smpl << 31) | ((unsigned)(dz + 8) << 16) | (unsigned)del);
    };
    // the kept offsets of a dz plane in (dy, dx) order, the order in which
    // the plain kernel sweeps a slab's candidates (dy outer, dx inner), so
    // that evaluating every candidate sums in the plain kernel's order: the
    // rank among the kept offsets by the key (dz, dy, dx), made once per lane
    // (every lane runs the shuffles; only lanes with a kept offset use the rank)
    const int key = has_k ? (kv / nw2) * nw2 + (kv % nw1) * nw1 + (kv / nw1) % nw1 : INT_MAX;
    int rank = 0;
    for (int q = 0; q < a.topk; ++q) rank += __shfl(key, q) < key;
    int* out = a.merged + item * kMergedStride;
    if (has_k) out[1 + rank + before_k] = pack(kv, 0u);
    if (has_s && !dup)
        out[1 + before_s + __popcll(smask & ((1ull << lane) - 1ull))] = pack(sv, 1u);
    if (lane == 0) {
        const int ns = __popcll(smask);
        out[0] = (__popcll(kmask) + ns) | (ns << 16);
    }
}

// Sums over the lattice samples of the lane's first output (z0, x = xb),
// for the error estimate of the tail term: w L and L (the sum of w is Ws).
// One output in eight of every tile is estimated, spread over all its rows
// and columns: enough to flag the tile, for two registers.
struct SampleMoments {
    float wl, l1;
};

// As tiled_step, also summing the CT of the evaluated candidates (Le) and
// the weights of the lattice samples (Ws); with ERR also the moments of the
// samples.
template <class C, int BX, int BZ, bool SAMPLE, bool ERR>
__device__ __forceinline__ void pruned_step(const float* __restrict__ lds, int base,
                                            const float (&tv)[BZ + 2][BX + 2],
                                            float (&P)[BZ][BX], float (&S)[BZ][BX],
                                            float (&Le)[BZ][BX], float (&Ws)[BZ][BX],
                                            SampleMoments& m) {
    constexpr int NBA = (BZ + 2 + C::PPB - 1) / C::PPB;
    constexpr int NBL = (BZ + C::PPB - 1) / C::PPB;
    int ba[NBA];
    int bl[NBL];
#pragma unroll
    for (int b = 0; b < NBA; ++b) ba[b] = b ? opaque(base + b * C::PPB * C::PS) : base;
#pragma unroll
    for (int b = 0; b < NBL; ++b) bl[b] = opaque(base + C::LOFF + b * C::PPB * C::PS);
    float ex[BZ + 2][BX];
#pragma unroll
    for (int pz = 0; pz < BZ + 2; ++pz) {
        float d[BX + 2];
        const int pp = ba[pz / C::PPB] + (pz % C::PPB) * C::PS;
#pragma unroll
        for (int i = 0; i < BX + 2; ++i) {
            const float df = tv[pz][i] - lds[pp + i * C::CW];
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
            const float l0 = lds[bl[j / C::PPB] + (j % C::PPB) * C::PS + (i + 1) * C::CW];
            const float l1 =
                lds[bl[(j + 1) / C::PPB] + ((j + 1) % C::PPB) * C::PS + (i + 1) * C::CW];
            P[j][i] = fmaf(w0, l0, P[j][i]);
            S[j][i] += w0;
            P[j + 1][i] = fmaf(w1, l1, P[j + 1][i]);
            S[j + 1][i] += w1;
            Le[j][i] += l0;
            Le[j + 1][i] += l1;
            if (SAMPLE) {
                Ws[j][i] += w0;
                Ws[j + 1][i] += w1;
            }
            if (SAMPLE && ERR && j == 0 && i == 0) {
                m.wl = fmaf(w0, l0, m.wl);
                m.l1 += l0;
            }
        }
    }
}

template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) void pruned_kernel(PrunedArgs pa) {
    using C = TiledCfg<W, BX, BZ, NW, HW>;
    extern __shared__ float lds[];
    // the offset list: per dz plane (ascending), the kept offsets, then the
    // lattice samples that are not kept; each entry's dz and LDS delta
    __shared__ int odz[kMaxOffsets];
    __shared__ int odel[kMaxOffsets];
    __shared__ int is_sample[kMaxOffsets];
    const TiledArgs& args = pa.t;
    const int item = xcd_contiguous(blockIdx.x, args.n_items);
    if (pa.flag[item]) return;            // full search handles it (uniform)
    const int at = item / args.n_tiles;
    int t = item - at * args.n_tiles;
    const int ty = t % args.tiles_y;
    t /= args.tiles_y;
    const int tx = t % args.tiles_x;
    const int tz = t / args.tiles_x;

    const int* __restrict__ mo = pa.merged + (long long)item * kMergedStride;
    const int hdr = mo[0];
    const int n_offs = hdr & 0xffff, n_samp_eff = hdr >> 16;
    if ((int)threadIdx.x < n_offs) {   // n_offs <= kMaxOffsets < NW * 64
        const unsigned e = (unsigned)mo[1 + threadIdx.x];
        odz[threadIdx.x] = (int)((e >> 16) & 15u) - 8;
        odel[threadIdx.x] = (int)(e & 0xffffu);
        is_sample[threadIdx.x] = (int)(e >> 31);
    }
    __syncthreads();

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
    float P[BZ][BX], S[BZ][BX], Le[BZ][BX], Ws[BZ][BX];
    SampleMoments mom{0.f, 0.f};
#pragma unroll
    for (int j = 0; j < BZ; ++j)
#pragma unroll
        for (int i = 0; i < BX; ++i) P[j][i] = S[j][i] = Le[j][i] = Ws[j][i] = 0.f;
    const bool err = pa.max_err > 0.f;   // uniform over the grid

    const int lane_base = sg * BX * C::CW + ly;
    const int xs = x0 - 1 - HW;
    const int ys = y0 - 1 - HW;
    DirectPlan<C, NW> dp;
    dp.init(xs, ys, nx, ny, wv, lane);
    const float* __restrict__ A = args.atlas + (long long)at * args.N;
    const float* __restrict__ L = args.label + (long long)at * args.N;
    const int m = n_offs;
    int k = 0;
    while (k < m) {
        const int dz = odz[k];
        __syncthreads();
        stage_slab_direct<C, NW>(lds, A, L, args.scale, z0 + dz, nz, plane, wv, dp);
        for (; k < m && odz[k] == dz && !is_sample[k]; ++k)
            pruned_step<C, BX, BZ, false, false>(lds, lane_base + odel[k], tv, P, S, Le, Ws, mom);
        if (err) {
            for (; k < m && odz[k] == dz; ++k)
                pruned_step<C, BX, BZ, true, true>(lds, lane_base + odel[k], tv, P, S, Le, Ws,
                                                   mom);
        } else {
            for (; k < m && odz[k] == dz; ++k)
                pruned_step<C, BX, BZ, true, false>(lds, lane_base + odel[k], tv, P, S, Le, Ws,
                                                    mom);
        }
    }

    // tail: the candidates not evaluated, at the mean sample weight
    int over = 0;
    if (ly >= 1 && ly <= W - 2 && y < ny - r) {
        const int ns = n_samp_eff;
        const float n_rest = (float)(pa.ncand - m);
        float* __restrict__ wsP = args.ws + (long long)(2 * at) * args.N;
        float* __restrict__ wsS = wsP + args.N;
        const float* __restrict__ bL = pa.boxL + (long long)at * args.N;
#pragma unroll
        for (int j = 0; j < BZ; ++j)
#pragma unroll
            for (int i = 0; i < BX; ++i) {
                const int z = z0 + j, x = xb + i;
                if (z < nz - r && x < nx - r) {
                    const long long idx = z * plane + (long long)x * ny + y;
                    const float wbar = ns > 0 ? Ws[j][i] / (float)ns : 0.f;
                    const float S_tail = wbar * n_rest;
                    const float S_all = fmaf(wbar, n_rest, S[j][i]);
                    wsP[idx] = fmaf(wbar, bL[idx] - Le[j][i], P[j][i]);
                    wsS[idx] = S_all;
                    if (err && ns > 1 && j == 0 && i == 0) {
                        // the tail stands for its candidates' weighted CT mean
                        // by their plain mean: it misses by cov(w, L) / mean(w)
                        // over them, estimated from the samples, times the
                        // tail's share of the weight
                        const float share = S_all > 0.f ? S_tail / S_all : 1.f;
                        const float cov_ratio =
                            Ws[0][0] > 0.f ? fabsf(fmaf(-wbar, mom.l1, mom.wl)) / Ws[0][0] : 0.f;
                        const float e = share * cov_ratio;
                        // an exhausted weight sum cannot be estimated at all
                        over |= !(e <= pa.max_err) || !(S_all > 0.f);
                    }
                }
            }
    }
    if (err && __syncthreads_or(over) && threadIdx.x == 0) pa.flag[item] = 2;
}

using PrunedFn = void (*)(PrunedArgs);

struct PrunedEntry {
    int shape_id;
    PrunedFn fn[kMaxTiledHalfWindow];
};

template <int W, int BX, int BZ, int NW>
PrunedEntry make_pruned(int shape_id) {
    return PrunedEntry{shape_id,
                       {&pruned_kernel<W, BX, BZ, NW, 1>, &pruned_kernel<W, BX, BZ, NW, 2>,
                        &pruned_kernel<W, BX, BZ, NW, 3>, &pruned_kernel<W, BX, BZ, NW, 4>,
                        &pruned_kernel<W, BX, BZ, NW, 5>, &pruned_kernel<W, BX, BZ, NW, 6>,
                        &pruned_kernel<W, BX, BZ, NW, 7>}};
}

static const std::vector<PrunedEntry>& pruned_entries() {
    static const std::vector<PrunedEntry> e = {
        make_pruned<32, 4, 2, 4>(8),
        make_pruned<32, 2, 4, 8>(13),
    };
    return e;
}

static const PrunedEntry* find_pruned(int shape_id) {
    for (const auto& e : pruned_entries())
        if (e.shape_id == shape_id) return &e;
    return nullptr;
}

}  // namespace pct

using namespace pct;

extern "C" int pct_mlp_init(float* params, uint64_t seed, float mu, float inv_sd) {
    if (!params || !(inv_sd > 0.f)) return PCT_ERR_INVALID;
    std::memset(params, 0, sizeof(float) * PCT_MLP_NPARAMS);
    uint64_t st = seed ^ 0x5DEECE66DULL;
    auto normal = [&]() {
        const double u1 = ((double)(splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int k = 0; k < 30; ++k)
        for (int m = 0; m < 64; ++m) params[kOffW1 + k * 64 + m] = (float)(normal() * std::sqrt(2.0 / 30));
    for (int k = 0; k < 64; ++k)
        for (int m = 0; m < 64; ++m) params[kOffW2 + k * 64 + m] = (float)(normal() * std::sqrt(2.0 / 64));
    for (int k = 0; k < 64; ++k)
        for (int m = 0; m < 16; ++m) params[kOffW3 + k * 16 + m] = (float)(normal() * std::sqrt(1.0 / 64));
    params[kOffMu] = mu;
    params[kOffSd] = inv_sd;
    return PCT_OK;
}

extern "C" int pct_embed_dev(const float* params, const float* d_vol, int nx, int ny, int nz,
                             uint16_t* d_emb, pct_stream_t stream) {
    if (!params || !d_vol || !d_emb || nx < 1 || ny < 1 || nz < 1) return PCT_ERR_INVALID;
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const void* model = nullptr;
    rc = device_model(dev, params, s, &model);
    if (rc) return rc;
    return launch_embed(params, model, d_vol, nx, ny, nz, 1,
                        reinterpret_cast<unsigned short*>(d_emb), s);
}

extern "C" int pct_fuse_pruned_dev(const pct_problem* pb, const float* params,
                                   const pct_pruned_opts* opts, const float* d_img,
                                   const float* d_atlas, const float* d_label, float* d_out,
                                   double* fallback_frac, pct_stream_t stream) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (!params || !opts || !d_img || !d_atlas || !d_label || !d_out) return PCT_ERR_INVALID;
    if (opts->topk < 1 || opts->topk > kMaxTopk || opts->samples < 0) return PCT_ERR_INVALID;
    const int hw = (pb->nh - 1) / 2;
    if (hw < 1 || hw > kMaxTiledHalfWindow) return PCT_ERR_UNSUPPORTED;
    // the direct LDS staging keeps in-plane source offsets in 32 bits
    if ((long long)pb->nx * pb->ny >= INT_MAX) return PCT_ERR_TOO_LARGE;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int shape_id = opts->shape_id >= 0 ? opts->shape_id : 13;
    const PrunedEntry* pe = find_pruned(shape_id);
    if (!pe) return PCT_ERR_UNSUPPORTED;
    TileGeometry g;
    if (!make_geometry(pb, shape_id, 1, &g)) return PCT_ERR_UNSUPPORTED;
    const TileShape& sh = shapes()[shape_id];
    const int n = pb->n_atlas;
    const long long N = (long long)pb->nx * pb->ny * pb->nz;
    const long long items = g.n_tiles * n;
    if (items >= INT_MAX) return PCT_ERR_TOO_LARGE;   // one workgroup per item, int item index
    const int ncand = (2 * hw + 1) * (2 * hw + 1) * (2 * hw + 1);
    const int topk = std::min(opts->topk, ncand);

    // workspace layout
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_part = off; off = align(off + sizeof(float) * 2 * (size_t)n * N);
    const size_t o_embT = off; off = align(off + 32 * (size_t)N);
    const size_t o_embA = off; off = align(off + 32 * (size_t)N * n);
    const size_t o_boxL = off; off = align(off + sizeof(float) * (size_t)N * n);
    const size_t o_sel = off; off = align(off + sizeof(int) * (size_t)items * topk);
    const size_t o_flag = off; off = align(off + (size_t)items);
    const size_t o_merged = off; off = align(off + sizeof(int) * (size_t)items * kMergedStride);
    void* ws = nullptr;
    rc = ensure_workspace(dev, off, &ws);
    if (rc) return rc;
    char* w = static_cast<char*>(ws);
    const void* model = nullptr;
    rc = device_model(dev, params, s, &model);
    if (rc) return rc;

    // the lattice samples are 3x3x3 offsets at 0 and +-ceil(2 hw / 3) (all
    // 27 offsets for NH = 3)
    const int nw = 2 * hw + 1;
    int samples[kSamples] = {0};
    int n_samples = 0;
    {
        const int q = std::max(1, (2 * hw + 2) / 3);
        const int lat[3] = {-q, 0, q};
        const int want = std::min(std::max(opts->samples, 0), kSamples);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                    if (n_samples < want)
                        samples[n_samples++] =
                            ((lat[a] + hw) * nw + (lat[b] + hw)) * nw + (lat[c] + hw);
        std::sort(samples, samples + n_samples);
    }
    int lattice_dz[3] = {0, 0, 0}, n_lattice_dz = 0;
    for (int i = 0; i < n_samples; ++i) {
        const int dz = samples[i] / (nw * nw) - hw;
        bool seen = false;
        for (int l = 0; l < n_lattice_dz; ++l) seen |= lattice_dz[l] == dz;
        if (!seen) lattice_dz[n_lattice_dz++] = dz;
    }
    // With at most as many slabs as lattice planes, kept offsets come from
    // the lattice planes only: the selection scores those dz alone, and the
    // atlas embeddings are needed only on the window planes they read.
    const bool lattice_only =
        opts->max_slabs > 0 && opts->max_slabs <= n_lattice_dz && n_lattice_dz < nw;
    unsigned score_bits = (1u << nw) - 1u;
    if (lattice_only) {
        score_bits = 0;
        for (int l = 0; l < n_lattice_dz; ++l) score_bits |= 1u << (lattice_dz[l] + hw);
    }
    std::vector<int> zlist;
    if (lattice_only && !opts->embed_all) {
        // window plane wz (0 .. 2 hw + 1) of tile row tz sits at z = pz0 - hw
        // + wz, clamped as the selection clamps it; plane wz is read by the
        // scored dz index wz (z = 0 probes) and wz - 1 (z = 1 probes)
        std::vector<char> need((size_t)pb->nz, 0);
        for (int tz = 0; tz < g.tiles_z; ++tz) {
            const int pz0 = probe_origin(tz, sh.bz, g.r, pb->nz, 2);
            for (int wz = 0; wz <= 2 * hw + 1; ++wz) {
                const bool rd = ((score_bits >> wz) & 1u) || (wz >= 1 && ((score_bits >> (wz - 1)) & 1u));
                if (rd) need[(size_t)std::min(std::max(pz0 - hw + wz, 0), pb->nz - 1)] = 1;
            }
        }
        for (int z = 0; z < pb->nz; ++z)
            if (need[(size_t)z]) zlist.push_back(z);
    }

    // the pruned kernel's arguments (problem, tiles, partials) and the input
    // it reads besides the selection and the volumes: the window sums of every
    // atlas CT, on the second stream beside steps 1 and 3.  The pruned kernel
    // copies the unscaled atlas MRI into LDS and scales it there.
    float* part = reinterpret_cast<float*>(w + o_part);
    PrunedArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    TileGeometry g1 = g;
    g1.group = 1;
    g1.n_groups = n;
    g1.n_items = items;
    pa.t = make_tiled_args(pb, g1, d_img, d_atlas, d_label, part);
    float* boxL = reinterpret_cast<float*>(w + o_boxL);
    std::unique_lock<std::mutex> side_lk(g_side.mu);   // until the join is enqueued on s
    rc = side_stream(dev);
    if (rc) return rc;
    hipStream_t ss = g_side.st[dev];
    if (hipEventRecord(g_side.fork[dev], s) != hipSuccess ||
        hipStreamWaitEvent(ss, g_side.fork[dev], 0) != hipSuccess)
        return PCT_ERR_HIP;
    {
        // 2. window sums of every atlas CT
        const long long blocks = (long long)n * ((pb->nz + kBoxRun - 1) / kBoxRun) *
                                 ((pb->nx + kBoxTX - 1) / kBoxTX) *
                                 ((pb->ny + kBoxTY - 1) / kBoxTY);
        hipLaunchKernelGGL(kBoxXyz[hw - 1], dim3((unsigned)blocks), dim3(kBoxThreads), 0, ss,
                           d_label, boxL, pb->nx, pb->ny, pb->nz, n);
        if (hipGetLastError() != hipSuccess || hipEventRecord(g_side.join[dev], ss) != hipSuccess)
            return PCT_ERR_HIP;
    }

    // 1. embeddings of the target and of every atlas MRI
    unsigned short* embT = reinterpret_cast<unsigned short*>(w + o_embT);
    unsigned short* embA = reinterpret_cast<unsigned short*>(w + o_embA);
    const int t_oy = sh.seg_width - 2, t_wx = sh.waves * (64 / sh.seg_width) * sh.bx;
    if (opts->embed_all) {
        rc = launch_embed(params, model, d_img, pb->nx, pb->ny, pb->nz, 1, embT, s);
        if (!rc) rc = launch_embed(params, model, d_atlas, pb->nx, pb->ny, pb->nz, n, embA, s);
    } else {
        rc = launch_embed_windows(params, model, d_img, 1, 0, pb, g, t_oy, t_wx, sh.bz, embT, s);
        const int* d_zlist = nullptr;
        if (!rc && !zlist.empty()) rc = device_zlist(dev, zlist, s, &d_zlist);
        if (!rc)
            rc = launch_embed_windows(params, model, d_atlas, n, hw, pb, g, t_oy, t_wx, sh.bz,
                                      embA, s, d_zlist, (int)zlist.size());
    }
    if (rc) return rc;

    // 3. selection
    SelectArgs sa;
    sa.embT = embT;
    sa.embA = embA;
    sa.sel = reinterpret_cast<int*>(w + o_sel);
    sa.flag = reinterpret_cast<unsigned char*>(w + o_flag);
    sa.N = N;
    sa.nx = pb->nx;
    sa.ny = pb->ny;
    sa.nz = pb->nz;
    sa.hw = hw;
    sa.r = g.r;
    sa.tiles_y = g.tiles_y;
    sa.tiles_x = g.tiles_x;
    sa.n_tiles = (int)g.n_tiles;
    sa.oy = sh.seg_width - 2;
    sa.wx = sh.waves * (64 / sh.seg_width) * sh.bx;
    sa.bz = sh.bz;
    sa.topk = topk;
    sa.margin = opts->margin;
    sa.max_slabs = opts->max_slabs;
    sa.n_lattice_dz = n_lattice_dz;
    for (int l = 0; l < 3; ++l) sa.lattice_dz[l] = lattice_dz[l];
    sa.score_bits = score_bits;
    // the allowed slabs the kernel would choose, where the scores cannot
    // change them: every slab, or the lattice's when they reach max_slabs
    {
        unsigned lat_bits = 0;
        for (int l = 0; l < n_lattice_dz; ++l) lat_bits |= 1u << (lattice_dz[l] + hw);
        if (opts->max_slabs <= 0 || opts->max_slabs >= nw)
            sa.fixed_allowed = (1u << nw) - 1u;
        else if (n_lattice_dz >= opts->max_slabs)
            sa.fixed_allowed = lat_bits;
        else
            sa.fixed_allowed = 0u;
    }
    if (ncand > kSelThreads * kSelPerThread) return PCT_ERR_UNSUPPORTED;
    // largest compacted candidate count: every slab, or the allowed ones
    const int slabs = (opts->max_slabs <= 0 || opts->max_slabs >= nw)
                          ? nw
                          : std::min(nw, std::max(opts->max_slabs, sa.n_lattice_dz));
    int p2 = 1;
    while (p2 < slabs * nw * nw) p2 <<= 1;
    // scores by offset, then either G of one window plane (while scoring)
    // or the sort arrays (after it)
    const int ring_stride = (((4 + 2 * hw) * (2 + 2 * hw) + 15) / 16) * 16;
    const size_t sel_lds = (size_t)((ncand + 63) & ~63) * 4 +
                           std::max((size_t)16 * ring_stride * 4, (size_t)p2 * 8);
    if (sel_lds > (size_t)kLdsPerCu) return PCT_ERR_UNSUPPORTED;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(select_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sel_lds);
    if (items > 0) {
        hipLaunchKernelGGL(select_kernel, dim3((unsigned)items), dim3(kSelThreads), sel_lds, s, sa);
        if (hipGetLastError() != hipSuccess) return PCT_ERR_HIP;
    }

    // 4. exact weighting of the kept candidates and the lattice samples,
    // after the second stream's work
    if (hipStreamWaitEvent(s, g_side.join[dev], 0) != hipSuccess) return PCT_ERR_HIP;
    side_lk.unlock();
    pa.sel = sa.sel;
    pa.flag = sa.flag;
    pa.boxL = boxL;
    pa.ncand = ncand;
    pa.max_err = opts->max_err > 0.f ? opts->max_err : 0.f;
    // the evaluated offsets of every item, merged once (one wave per item)
    {
        MergeArgs ma;
        ma.sel = sa.sel;
        ma.merged = reinterpret_cast<int*>(w + o_merged);
        ma.items = items;
        ma.topk = topk;
        ma.hw = hw;
        ma.cw = sh.seg_width + 2 * hw;   // LDS row length of the tiled slab (TiledCfg::CW)
        ma.n_samples = n_samples;
        std::memcpy(ma.samples, samples, sizeof(samples));
        if (items > 0) {
            hipLaunchKernelGGL(merge_offsets_kernel,
                               dim3((unsigned)((items + kMergeWaves - 1) / kMergeWaves)),
                               dim3(kMergeWaves * 64), 0, s, ma);
            if (hipGetLastError() != hipSuccess) return PCT_ERR_HIP;
        }
        pa.merged = ma.merged;
    }
    PrunedFn fn = pe->fn[hw - 1];
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                              hipFuncAttributeMaxDynamicSharedMemorySize, g.lds_bytes);
    if (items > 0) {
        void* kargs[] = {&pa};
        TimedLaunch timed(s);
        if (hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3((unsigned)items),
                            dim3(sh.waves * 64), kargs, (size_t)g.lds_bytes, s) != hipSuccess)
            return PCT_ERR_HIP;
    }

    // 5. full exact search for the flagged items, by the selection's margin
    // or the pruned kernel's error estimate (with neither, nothing is
    // flagged and the launch is skipped)
    if (opts->margin != 0.f || pa.max_err > 0.f) {
        TiledArgs fa = pa.t;
        fa.only = sa.flag;
        rc = launch_tiled_items(fa, shape_id, hw, s);
        if (rc) return rc;
    }

    // 6. sum over atlases, NaN rule, zero border
    rc = launch_finalize(pb, g1, part, d_out, s);
    if (rc) return rc;

    if (fallback_frac) {
        std::vector<unsigned char> fl((size_t)items);
        if (items > 0 && hipMemcpyAsync(fl.data(), sa.flag, (size_t)items, hipMemcpyDeviceToHost,
                                        s) != hipSuccess)
            return PCT_ERR_HIP;
        if (hipStreamSynchronize(s) != hipSuccess) return PCT_ERR_HIP;
        long long cnt = 0;
        for (unsigned char c : fl) cnt += c != 0;
        *fallback_frac = items > 0 ? (double)cnt / (double)items : 0.0;
    }
    return PCT_OK;
}
