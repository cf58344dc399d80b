// Post-fusion chain on the GPU: NaN fill by the median of the 3x3x3
// neighbourhood, rounding, and the HU rescale, with the reference's order
// of operations (src/CPU/NewPatchedBrainPureC.cpp:1020-1106 Median and
// RegularizationStep, 1221-1227 RoundMatrix, 1725-1739 the HU rescale).
//
// RegularizationStep sweeps the interior in z, x, y order and overwrites each
// NaN voxel in place, so a voxel sees the values filled earlier in the same
// sweep.  Only NaN voxels ever change, and a NaN voxel's median reads its
// neighbours: those before it in scan order as this sweep left them, those
// after it as the previous sweep did (a NaN one after it is still NaN when
// it is reached).  With the wavefront index t = 4z + 2x + y every neighbour
// before a voxel in scan order has a smaller t and every neighbour after it
// a larger one (the steepest neighbours differ by (0, 0, +-1), (0, +-1, -+1)
// and (+-1, -+1, -+1): t steps of 1, 1 and 1).  Filling the sweep's NaN
// voxels in increasing t, all voxels of one t at once, therefore gives the
// sequential sweep's values bit for bit.
//
// (1) The interior NaN voxels are listed (per-chunk counts, a scan, a
// compaction per chunk); (2) one workgroup runs the sweeps: a counting sort
// of the list by t, then the levels in increasing t with a barrier after
// each non-empty one, one voxel per thread (the median by stable ranks, the
// std::stable_sort of cpp:1035, signed zeros included).  Voxels whose median
// is still NaN form the next sweep's list; sweeps stop when no NaN is left
// or a sweep fills nothing.  Rounding and the rescale are elementwise.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "pct_internal.h"
#include "pct_tiled.h"

namespace pct {

// Interior NaN voxels of `arr` in scan order (z, x, y; y fastest) into
// list[0..*count), in three passes over chunks of kListChunk interior
// positions (one block each): count per chunk, exclusive scan of the counts,
// then an ordered compaction of every chunk at its offset.
constexpr int kListThreads = 256;
constexpr int kListChunk = 4096;

__device__ __forceinline__ int interior_id(long long t, int lo, int ix, int iy, int nx, int ny) {
    const int y = lo + (int)(t % iy);
    const int x = lo + (int)((t / iy) % ix);
    const int z = lo + (int)(t / ((long long)iy * ix));
    return (z * nx + x) * ny + y;
}

__global__ __launch_bounds__(kListThreads) void nan_count_kernel(const float* __restrict__ arr,
                                                                 int nx, int ny, int lo, int ix,
                                                                 int iy, long long total,
                                                                 int* __restrict__ bcount) {
    __shared__ int wsum[kListThreads / 64];
    const long long c0 = (long long)blockIdx.x * kListChunk;
    int cnt = 0;
    for (int k = threadIdx.x; k < kListChunk; k += kListThreads) {
        const long long t = c0 + k;
        if (t < total && isnan(arr[interior_id(t, lo, ix, iy, nx, ny)])) ++cnt;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < kListThreads / 64; ++w) tot += wsum[w];
        bcount[blockIdx.x] = tot;
    }
}

// Exclusive scan of nb block counts into boff; the total into *count.
__global__ __launch_bounds__(1024) void nan_scan_kernel(const int* __restrict__ bcount, int nb,
                                                        int* __restrict__ boff,
                                                        int* __restrict__ count) {
    __shared__ int wtot[16];
    const int seg = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * seg, b1 = min(nb, b0 + seg);
    int own = 0;
    for (int b = b0; b < b1; ++b) own += bcount[b];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int before = 0;
    for (int j = 0; j < w; ++j) before += wtot[j];
    int run = before + incl - own;   // exclusive prefix of this thread's segment
    for (int b = b0; b < b1; ++b) {
        boff[b] = run;
        run += bcount[b];
    }
    if (threadIdx.x == 1023) {
        int tot = 0;
        for (int j = 0; j < 16; ++j) tot += wtot[j];
        *count = tot;
    }
}

__global__ __launch_bounds__(kListThreads) void nan_write_kernel(const float* __restrict__ arr,
                                                                 int nx, int ny, int lo, int ix,
                                                                 int iy, long long total,
                                                                 const int* __restrict__ boff,
                                                                 int* __restrict__ list) {
    __shared__ int wcnt[kListThreads / 64];
    const long long c0 = (long long)blockIdx.x * kListChunk;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int base = boff[blockIdx.x];
    for (int r = 0; r < kListChunk; r += kListThreads) {
        const long long t = c0 + r + threadIdx.x;
        int id = 0;
        bool isn = false;
        if (t < total) {
            id = interior_id(t, lo, ix, iy, nx, ny);
            isn = isnan(arr[id]);
        }
        const unsigned long long bal = __ballot(isn);
        if (lane == 0) wcnt[w] = __popcll(bal);
        __syncthreads();
        int off = base, tot = 0;
#pragma unroll
        for (int j = 0; j < kListThreads / 64; ++j) {
            if (j < w) off += wcnt[j];
            tot += wcnt[j];
        }
        if (isn) list[off + __popcll(bal & ((1ull << lane) - 1ull))] = id;
        __syncthreads();   // wcnt is rewritten next round
        base += tot;
    }
}

constexpr int kFillThreads = 1024;

// Comparators of Batcher's merge-exchange sorting network for 27 inputs
// (Knuth, TAOCP vol. 3, 5.2.2 Algorithm M; valid for any input count):
// 155 (min, max) pairs, built at compile time.
struct SortNet27 {
    int n = 0;
    unsigned char lo[160] = {}, hi[160] = {};
};
constexpr SortNet27 merge_exchange27() {
    SortNet27 L{};
    constexpr int N = 27;
    int t = 0;
    while ((1 << t) < N) ++t;
    for (int p = 1 << (t - 1); p > 0; p >>= 1) {
        int q = 1 << (t - 1), r = 0, d = p;
        for (;;) {
            for (int i = 0; i < N - d; ++i)
                if ((i & p) == r) {
                    L.lo[L.n] = (unsigned char)i;
                    L.hi[L.n] = (unsigned char)(i + d);
                    ++L.n;
                }
            if (q == p) break;
            d = q - p;
            q >>= 1;
            r = p;
        }
    }
    return L;
}
constexpr SortNet27 kSortNet27 = merge_exchange27();
static_assert(kSortNet27.n == 155, "merge exchange for 27 inputs");

// Order-preserving integer key of a float that is not NaN (an involution:
// applied to a key it gives the float's bits back).  -0 sorts before +0.
__device__ __forceinline__ int float_key(float f) {
    const int u = __float_as_int(f);
    return u ^ (int)((unsigned)(u >> 31) >> 1);
}

// Median of the non-NaN values among `cnt` of v[27] by stable ranks: the
// std::stable_sort of cpp:1035, equal values in their gathered order.
__device__ __forceinline__ float median_by_ranks(const float (&v)[27], int cnt) {
    const int hi = cnt / 2, lo_rank = cnt / 2 - 1;
    float a = 0.f, b = 0.f;
#pragma unroll
    for (int i = 0; i < 27; ++i) {
        if (isnan(v[i])) continue;
        int rank = 0;
#pragma unroll
        for (int j = 0; j < 27; ++j)
            rank += (!isnan(v[j]) && (v[j] < v[i] || (v[j] == v[i] && j < i))) ? 1 : 0;
        if (rank == hi) b = v[i];
        if (rank == lo_rank) a = v[i];
    }
    if (cnt % 2 == 0) return (float)(((double)a + (double)b) / 2);
    return b;
}

// Median of the 3x3x3 neighbourhood of voxel `id` over its non-NaN values
// (cpp:1020-1043): the middle value, or the float64 mean of the two middle
// values, of their stable sort; NaN when all 27 are NaN.  Values that compare
// equal have the same bits except -0 and +0, so unless both signed zeros
// occur, any sort of the values gives the stable sort's bits: a sorting
// network of integer min/max pairs over order-preserving keys (NaN keyed
// last), 155 pairs against 27 x 27 rank comparisons.  With both signed zeros
// the ranks decide.
__device__ __forceinline__ float neighbourhood_median(const float* arr, int id, int pl, int ny) {
    float v[27];
#pragma unroll
    for (int k = 0; k < 27; ++k)
        v[k] = __hip_atomic_load(arr + id + (k / 9 - 1) * pl + ((k / 3) % 3 - 1) * ny + (k % 3 - 1),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    int cnt = 0;
    bool neg_zero = false, pos_zero = false;
    int key[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const bool nan = isnan(v[k]);
        cnt += nan ? 0 : 1;
        const int u = __float_as_int(v[k]);
        neg_zero |= u == (int)0x80000000;
        pos_zero |= u == 0;
        key[k] = nan ? INT_MAX : float_key(v[k]);
    }
    if (cnt == 0) return __builtin_nanf("");
    if (neg_zero && pos_zero) return median_by_ranks(v, cnt);
#pragma unroll
    for (int c = 0; c < kSortNet27.n; ++c) {
        const int i = kSortNet27.lo[c], j = kSortNet27.hi[c];
        const int mn = min(key[i], key[j]);
        key[j] = max(key[i], key[j]);
        key[i] = mn;
    }
    const int hi = cnt / 2, lo_rank = cnt / 2 - 1;
    int ka = 0, kb = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        kb = k == hi ? key[k] : kb;
        ka = k == lo_rank ? key[k] : ka;
    }
    const float a = __int_as_float(float_key(__int_as_float(ka))), b = __int_as_float(float_key(__int_as_float(kb)));
    if (cnt % 2 == 0) return (float)(((double)a + (double)b) / 2);
    return b;
}

// The sweeps of RegularizationStep over the listed voxels, in one workgroup.
// hist has t_count + 1 ints; sorted and the lists hold `interior` ints.
// Every thread that reads or writes these buffers during the fill belongs to
// this one workgroup (the inputs come from earlier launches on the stream),
// so values and list entries go through workgroup-scope atomics and every
// level ends with a workgroup barrier, whose workgroup-scope release and
// acquire let a level see the ones before it.  Agent-scope fences would
// write back and invalidate the L2 at every level.  The level histogram is
// built by atomics performed in L2 and read back with agent-scope loads.
__global__ __launch_bounds__(kFillThreads) void nan_fill_levels_kernel(
    float* __restrict__ arr, int* list_a, int* list_b, int* sorted, int* hist,
    const int* count0, int nx, int ny, int t_count, int* sweeps_out) {
    __shared__ int s_n, s_next, s_filled;
    __shared__ int s_end[kFillThreads];   // level ends of one chunk of levels
    const int pl = nx * ny;
    int* cur = list_a;
    int* nxt = list_b;
    if (threadIdx.x == 0) s_n = __hip_atomic_load(count0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    int n = s_n;
    int sweeps = 0;
    while (n > 0) {
        ++sweeps;
        // counting sort of the list by t = 4 z + 2 x + y
        for (int t = threadIdx.x; t <= t_count; t += kFillThreads) hist[t] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kFillThreads) {
            const int id = cur[i];
            const int z = id / pl, x = (id - z * pl) / ny, y = id - z * pl - x * ny;
            atomicAdd(hist + 4 * z + 2 * x + y, 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // exclusive prefix of the counts; hist[t] becomes the level start.
            // The counts come from atomics performed in L2: read them with
            // agent-scope loads, which do not use the CU's L1, 16 at a time.
            int run = 0;
            for (int t0 = 0; t0 <= t_count; t0 += 16) {
                int c[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    c[k] = t0 + k <= t_count ? __hip_atomic_load(hist + t0 + k, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT)
                                             : 0;
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (t0 + k <= t_count) {
                        hist[t0 + k] = run;
                        run += c[k];
                    }
            }
            s_next = 0;
            s_filled = 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kFillThreads) {
            const int id = cur[i];
            const int z = id / pl, x = (id - z * pl) / ny, y = id - z * pl - x * ny;
            sorted[atomicAdd(hist + 4 * z + 2 * x + y, 1)] = id;
        }
        __syncthreads();
        // hist[t] is now the end of level t; levels in increasing t, their
        // ends read into LDS a chunk at a time
        int begin = 0;
        for (int t0 = 0; t0 < t_count; t0 += kFillThreads) {
            if (t0 + (int)threadIdx.x < t_count)
                s_end[threadIdx.x] = __hip_atomic_load(hist + t0 + threadIdx.x, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);   // L2 atomics
            __syncthreads();
            const int tn = min(kFillThreads, t_count - t0);
            for (int k = 0; k < tn; ++k) {
                const int end = s_end[k];
                if (end == begin) continue;   // uniform: an empty level
                for (int i = begin + (int)threadIdx.x; i < end; i += kFillThreads) {
                    const int id = __hip_atomic_load(sorted + i, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float res = neighbourhood_median(arr, id, pl, ny);
                    __hip_atomic_store(arr + id, res, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (isnan(res))
                        __hip_atomic_store(nxt + atomicAdd(&s_next, 1), id, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    else
                        atomicAdd(&s_filled, 1);
                }
                __syncthreads();
                begin = end;
            }
            __syncthreads();   // s_end is rewritten for the next chunk
        }
        const int m = s_next, filled = s_filled;
        __syncthreads();   // every thread read the counters
        if (m == 0 || filled == 0) break;
        int* tmp = cur;
        cur = nxt;
        nxt = tmp;
        n = m;
    }
    if (threadIdx.x == 0 && sweeps_out) *sweeps_out = sweeps;
}

// RoundMatrix (cpp:1221-1227) then round(4095 * (v / 255)) - 1024 in float
// (cpp:1732-1738); `round` rounds halves away from zero on both sides.
__global__ void round_hu_kernel(float* __restrict__ arr, long long n, int to_hu) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float v = roundf(arr[i]);
        if (to_hu) v = roundf(4095.0f * (v / 255.0f)) - 1024.0f;
        arr[i] = v;
    }
}

static int post_checks(int nh, int nx, int ny, int nz, const float* d_arr) {
    if (!d_arr || nh < 1 || nx < 1 || ny < 1 || nz < 1) return PCT_ERR_INVALID;
    if ((long long)nx * ny * nz >= INT_MAX) return PCT_ERR_TOO_LARGE;
    return PCT_OK;
}

// Interior voxel count of the post chain's NaN fill.
static long long fill_interior(int nh, int nx, int ny, int nz) {
    const int lo = (nh - 1) / 2 + 1;
    const long long iz = nz - 2 * lo, ix = nx - 2 * lo, iy = ny - 2 * lo;
    return (iz > 0 && ix > 0 && iy > 0) ? iz * ix * iy : 0;
}

// Wavefront levels of a volume: t = 4 z + 2 x + y ranges over [0, t_count).
static long long fill_levels(int nx, int ny, int nz) {
    return 4LL * (nz - 1) + 2LL * (nx - 1) + (ny - 1) + 1;
}

// Workspace of the NaN fill: a counter, three voxel lists (two sweeps' and
// the sorted one), per-chunk counts and offsets, and the level histogram.
static size_t fill_ws_bytes(long long interior, long long levels) {
    const size_t list_bytes = ((size_t)interior * 4 + 255) & ~(size_t)255;
    const long long nb = (interior + kListChunk - 1) / kListChunk;
    const size_t block_bytes = ((size_t)nb * 4 + 255) & ~(size_t)255;
    const size_t hist_bytes = ((size_t)(levels + 1) * 4 + 255) & ~(size_t)255;
    return 256 + 3 * list_bytes + 2 * block_bytes + hist_bytes;
}

static int enqueue_fill(int nh, int nx, int ny, int nz, float* d_arr, void* ws, hipStream_t s) {
    const int lo = (nh - 1) / 2 + 1;
    const long long ix = nx - 2 * lo, iy = ny - 2 * lo;
    const long long interior = fill_interior(nh, nx, ny, nz);
    if (interior == 0) return PCT_OK;
    const size_t list_bytes = ((size_t)interior * 4 + 255) & ~(size_t)255;
    const long long nb = (interior + kListChunk - 1) / kListChunk;
    const size_t block_bytes = ((size_t)nb * 4 + 255) & ~(size_t)255;
    char* w = static_cast<char*>(ws);
    int* count = reinterpret_cast<int*>(w);
    int* la = reinterpret_cast<int*>(w + 256);
    int* lb = reinterpret_cast<int*>(w + 256 + list_bytes);
    int* srt = reinterpret_cast<int*>(w + 256 + 2 * list_bytes);
    int* bcount = reinterpret_cast<int*>(w + 256 + 3 * list_bytes);
    int* boff = reinterpret_cast<int*>(w + 256 + 3 * list_bytes + block_bytes);
    int* hist = reinterpret_cast<int*>(w + 256 + 3 * list_bytes + 2 * block_bytes);
    const long long levels = fill_levels(nx, ny, nz);
    const int ixi = (int)ix, iyi = (int)iy;
    hipLaunchKernelGGL(nan_count_kernel, dim3((unsigned)nb), dim3(kListThreads), 0, s, d_arr, nx,
                       ny, lo, ixi, iyi, interior, bcount);
    hipLaunchKernelGGL(nan_scan_kernel, dim3(1), dim3(1024), 0, s, bcount, (int)nb, boff, count);
    hipLaunchKernelGGL(nan_write_kernel, dim3((unsigned)nb), dim3(kListThreads), 0, s, d_arr, nx,
                       ny, lo, ixi, iyi, interior, boff, la);
    hipLaunchKernelGGL(nan_fill_levels_kernel, dim3(1), dim3(kFillThreads), 0, s, d_arr, la, lb,
                       srt, hist, count, nx, ny, (int)levels, (int*)nullptr);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

// The fill on the library's per-device workspace.
static int enqueue_fill_lib(int nh, int nx, int ny, int nz, float* d_arr, hipStream_t s) {
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    const long long interior = fill_interior(nh, nx, ny, nz);
    if (interior == 0) return PCT_OK;
    void* ws = nullptr;
    rc = ensure_workspace(dev, fill_ws_bytes(interior, fill_levels(nx, ny, nz)), &ws);
    if (rc) return rc;
    return enqueue_fill(nh, nx, ny, nz, d_arr, ws, s);
}

static int enqueue_round_hu(int nx, int ny, int nz, float* d_arr, hipStream_t s) {
    const long long n = (long long)nx * ny * nz;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(round_hu_kernel, dim3(blocks), dim3(256), 0, s, d_arr, n, 1);
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

size_t postprocess_ws_bytes(int nh, int nx, int ny, int nz) {
    return fill_ws_bytes(fill_interior(nh, nx, ny, nz), fill_levels(nx, ny, nz));
}

int postprocess_dev_ws(int nh, int nx, int ny, int nz, float* d_arr, void* ws, hipStream_t s) {
    int rc = post_checks(nh, nx, ny, nz, d_arr);
    if (rc) return rc;
    rc = enqueue_fill(nh, nx, ny, nz, d_arr, ws, s);
    if (rc) return rc;
    return enqueue_round_hu(nx, ny, nz, d_arr, s);
}

}  // namespace pct

using namespace pct;

extern "C" int pct_regularize_dev(int nh, int nx, int ny, int nz, float* d_arr,
                                  pct_stream_t stream) {
    const int rc = post_checks(nh, nx, ny, nz, d_arr);
    if (rc) return rc;
    return enqueue_fill_lib(nh, nx, ny, nz, d_arr, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_postprocess_dev(int nh, int nx, int ny, int nz, float* d_arr,
                                   pct_stream_t stream) {
    int rc = post_checks(nh, nx, ny, nz, d_arr);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    rc = enqueue_fill_lib(nh, nx, ny, nz, d_arr, s);
    if (rc) return rc;
    return enqueue_round_hu(nx, ny, nz, d_arr, s);
}
