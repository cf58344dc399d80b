// Batched fusion: many problems (subjects) of uneven volume size, atlas
// count and search window in one graph-captured launch of persistent kernels.
//
// The work items of the subjects - (tile, atlas group) pairs of the tiled
// kernel - are grouped by search-window class (half window HW).  Every class
// has its own tile shape and atlas group and one persistent kernel, fully
// specialised on HW (the register allocation of that window, not the union
// over all windows), with as many workgroups as the GPU holds at once.
//
// Queues follow the XCD split of the single-problem kernel: each subject's
// items are cut into eight contiguous parts, and chunk x of a class queue is
// the x-th part of every subject of the class, longest items first.  A
// workgroup starts on the chunk of its own XCD (workgroups are dispatched
// round-robin over the eight XCDs), so tiles that share halos share an L2,
// and moves on to the other chunks once its own is drained.  Each chunk has
// a device-scope counter; every workgroup leaves after it has found all
// eight chunks drained.  A subject's result is bit-identical to its
// individual run with the same tile shape and atlas group.
//
// The class kernels run either concurrently, as branches of the graph, so
// the tail of one class overlaps the next, or one after another; the
// autotuner measures both.  The counter reset, the direct kernel for windows
// without a tiled specialisation and the per-subject finalize passes
// complete the graph, which is replayed while the subject descriptors stay
// the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "pct_internal.h"
#include "pct_tiled.h"

namespace pct {

constexpr int kChunks = 8;                  // one queue chunk per XCD
constexpr size_t kQueueBytes = 512;         // counters at the workspace start
static_assert(kMaxTiledHalfWindow * kChunks * 8 <= (int)kQueueBytes, "queue counters");
static_assert(PCT_BATCH_CLASSES == kMaxTiledHalfWindow + 1, "class index is the half window");

struct BatchArgs {
    TiledArgs sub[PCT_MAX_BATCH];            // subjects of one window class, in queue order
    long long n_items[PCT_MAX_BATCH];
    int n_sub;
    unsigned long long* counters;            // kChunks claim counters of the class
};
static_assert(sizeof(BatchArgs) <= 4096, "batch descriptor must fit the kernel arguments");

// Part x of n items split into eight contiguous parts (the split of
// xcd_contiguous): its length and first item.
__device__ __forceinline__ long long part_len(long long n, int x) {
    return (n >> 3) + ((n & 7) > x ? 1 : 0);
}
__device__ __forceinline__ long long part_start(long long n, int x) {
    const long long per = n >> 3, rem = n & 7;
    return x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
}

template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) void fuse_batched(BatchArgs args) {
    extern __shared__ float lds[];
    __shared__ long long claimed;
    const int home = blockIdx.x & (kChunks - 1);
    for (int c = 0; c < kChunks; ++c) {
        const int x = (home + c) & (kChunks - 1);
        for (;;) {
            if (threadIdx.x == 0) claimed = (long long)atomicAdd(args.counters + x, 1ULL);
            __syncthreads();
            long long pos = claimed;
            __syncthreads();
            int k = 0;
            for (; k < args.n_sub; ++k) {
                const long long len = part_len(args.n_items[k], x);
                if (pos < len) break;
                pos -= len;
            }
            if (k == args.n_sub) break;   // chunk x is drained: every later claim is too
            tiled_item<W, BX, BZ, NW, HW>(
                args.sub[k], (int)(part_start(args.n_items[k], x) + pos), lds);
        }
    }
}

// fuse_batched with the cross-term items (TiledArgs::le filled per subject).
// The item loop costs the compiler some 18 VGPRs over the one-item kernel;
// past 128 a SIMD holds three waves instead of four, which the LDS of the
// narrow shapes (37.6 KB at W 32, BX 2, BZ 2, HW 5) would allow, and the
// class runs a quarter slower.  Four waves per SIMD caps it at 128.
template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4)))
void fuse_batched_xt(BatchArgs args) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ long long claimed;
    const int home = blockIdx.x & (kChunks - 1);
    for (int c = 0; c < kChunks; ++c) {
        const int x = (home + c) & (kChunks - 1);
        for (;;) {
            if (threadIdx.x == 0) claimed = (long long)atomicAdd(args.counters + x, 1ULL);
            __syncthreads();
            long long pos = claimed;
            __syncthreads();
            int k = 0;
            for (; k < args.n_sub; ++k) {
                const long long len = part_len(args.n_items[k], x);
                if (pos < len) break;
                pos -= len;
            }
            if (k == args.n_sub) break;   // chunk x is drained: every later claim is too
            tiled_item_xt<W, BX, BZ, NW, HW>(
                args.sub[k], (int)(part_start(args.n_items[k], x) + pos), lds);
        }
    }
}

// Written as a kernel rather than a memset so that it can sit in a captured
// graph.
__global__ void reset_queues(unsigned long long* counters, int n) {
    if ((int)threadIdx.x < n) counters[threadIdx.x] = 0ULL;
}

using BatchedFn = void (*)(BatchArgs);

struct BatchKernel {
    int shape_id;                          // index into shapes()
    BatchedFn fn[kMaxTiledHalfWindow];     // by HW - 1
};

template <int W, int BX, int BZ, int NW>
BatchKernel make_batch_kernel(int shape_id) {
    return BatchKernel{shape_id,
                       {&fuse_batched<W, BX, BZ, NW, 1>, &fuse_batched<W, BX, BZ, NW, 2>,
                        &fuse_batched<W, BX, BZ, NW, 3>, &fuse_batched<W, BX, BZ, NW, 4>,
                        &fuse_batched<W, BX, BZ, NW, 5>, &fuse_batched<W, BX, BZ, NW, 6>,
                        &fuse_batched<W, BX, BZ, NW, 7>}};
}

template <int W, int BX, int BZ, int NW>
BatchKernel make_batch_kernel_xt(int shape_id) {
    return BatchKernel{shape_id,
                       {&fuse_batched_xt<W, BX, BZ, NW, 1>, &fuse_batched_xt<W, BX, BZ, NW, 2>,
                        &fuse_batched_xt<W, BX, BZ, NW, 3>, &fuse_batched_xt<W, BX, BZ, NW, 4>,
                        &fuse_batched_xt<W, BX, BZ, NW, 5>, &fuse_batched_xt<W, BX, BZ, NW, 6>,
                        &fuse_batched_xt<W, BX, BZ, NW, 7>}};
}

// Tile shapes with persistent instantiations (ids of shapes()).
static const std::vector<BatchKernel>& batch_kernels() {
    static const std::vector<BatchKernel> k = {
        make_batch_kernel<32, 2, 4, 4>(7),
        make_batch_kernel<32, 4, 2, 4>(8),
        make_batch_kernel<64, 2, 4, 4>(9),
        make_batch_kernel<32, 2, 2, 8>(10),
        make_batch_kernel<32, 2, 2, 4>(11),
        make_batch_kernel<32, 2, 4, 8>(13),
        make_batch_kernel_xt<32, 2, 2, 4>(16),
        make_batch_kernel_xt<32, 2, 2, 8>(17),
        make_batch_kernel_xt<64, 2, 2, 4>(19),
    };
    return k;
}

static const BatchKernel* find_batch_kernel(int shape_id) {
    for (const auto& k : batch_kernels())
        if (k.shape_id == shape_id) return &k;
    return nullptr;
}

static bool tiled_window(const pct_problem& pb) {
    const int hw = (pb.nh - 1) / 2;
    return hw >= 1 && hw <= kMaxTiledHalfWindow;
}

// A tiled subject has work items when its interior (border r = HW + 1 on
// every side) is not empty; that does not depend on the tile shape.
static bool has_items(const pct_problem& pb) {
    if (!tiled_window(pb)) return false;
    const int r = (pb.nh - 1) / 2 + 1;
    return pb.nx > 2 * r && pb.ny > 2 * r && pb.nz > 2 * r;
}

static int class_of(const pct_problem& pb) { return (pb.nh - 1) / 2; }

// Layout of a batch in the workspace and in the queues for one configuration.
struct BatchPlan {
    pct_batch_choice choice;
    int n = 0;
    std::vector<TileGeometry> geo;            // per subject (input order)
    std::vector<size_t> ws_off;               // per subject, bytes from the workspace start
    std::vector<std::vector<int>> classes;    // subjects with items by window class
    std::vector<int> class_hw;
    std::vector<int> class_lds;
    size_t ws_bytes = 0;
};

static bool plan_batch(const pct_subject* s, int n, const pct_batch_choice& ch, BatchPlan* p) {
    p->choice = ch;
    p->n = n;
    p->geo.assign(n, TileGeometry());
    p->ws_off.assign(n, 0);
    p->classes.clear();
    p->class_hw.clear();
    p->class_lds.clear();
    size_t off = kQueueBytes;
    std::vector<double> item_cost(n, 0.0);
    std::vector<int> by_hw[kMaxTiledHalfWindow + 1];
    for (int i = 0; i < n; ++i) {
        if (!tiled_window(s[i].pb)) continue;
        const int hw = class_of(s[i].pb);
        int sid = ch.shape_id[hw], grp = ch.atlas_group[hw];
        if (sid < 0) {
            // a class whose subjects have no items: the geometry only sizes
            // the finalize pass
            if (has_items(s[i].pb)) return false;
            sid = batch_kernels().front().shape_id;
            grp = 1;
        }
        if (!find_batch_kernel(sid) || grp < 1) return false;
        TileGeometry g;
        if (!make_geometry(&s[i].pb, sid, std::min(grp, s[i].pb.n_atlas), &g)) return false;
        p->geo[i] = g;
        p->ws_off[i] = off;
        off += tiled_ws_bytes(&s[i].pb, g.n_groups, sid);
        item_cost[i] = g.group;
        if (g.n_items > 0) by_hw[hw].push_back(i);
    }
    // classes with the widest windows first; inside a class, longest items first
    for (int hw = kMaxTiledHalfWindow; hw >= 1; --hw) {
        if (by_hw[hw].empty()) continue;
        std::stable_sort(by_hw[hw].begin(), by_hw[hw].end(),
                         [&](int a, int b) { return item_cost[a] > item_cost[b]; });
        int lds = 0;
        for (int i : by_hw[hw]) lds = std::max(lds, p->geo[i].lds_bytes);
        p->classes.push_back(by_hw[hw]);
        p->class_hw.push_back(hw);
        p->class_lds.push_back(lds);
    }
    p->ws_bytes = off;
    return true;
}

static int validate_batch(const pct_subject* s, int n) {
    if (!s || n < 1) return PCT_ERR_INVALID;
    if (n > PCT_MAX_BATCH) return PCT_ERR_TOO_LARGE;
    for (int i = 0; i < n; ++i) {
        const int rc = validate_problem(&s[i].pb);
        if (rc) return rc;
        if (!s[i].d_img || !s[i].d_atlas || !s[i].d_atlas_label_norm || !s[i].d_out)
            return PCT_ERR_INVALID;
    }
    return PCT_OK;
}

// Workgroups of a persistent grid: every CU filled to the kernel's occupancy.
static int persistent_grid(int dev, int shape_id, BatchedFn fn, int lds_bytes) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(fn), shapes()[shape_id].waves * 64,
            (size_t)lds_bytes) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    return prop.multiProcessorCount * per_cu;
}

// Enqueues one batched run on `stream`, which is being captured.  With
// concurrent classes the class kernels fork onto `side` streams and join
// back before the finalize passes; serial classes stay on `stream`.
static int enqueue_batch(const pct_subject* s, const BatchPlan& p, const std::vector<int>& grids,
                         void* ws, hipStream_t stream, std::vector<hipStream_t>& side,
                         std::vector<hipEvent_t>& events) {
    char* base = static_cast<char*>(ws);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(base);
    const int nc = (int)p.classes.size();
    const bool concurrent = !p.choice.serial && nc > 1;
    hipLaunchKernelGGL(reset_queues, dim3(1), dim3(64), 0, stream, counters, nc * kChunks);
    if (hipGetLastError() != hipSuccess) return PCT_ERR_HIP;
    // cross-term pairs of the atlases of every subject of a cross-term class,
    // before the class kernels (and the fork of concurrent classes)
    for (int c = 0; c < nc; ++c) {
        if (!is_xterm_shape(p.choice.shape_id[p.class_hw[c]])) continue;
        for (int i : p.classes[c]) {
            const TiledArgs t = make_tiled_args(&s[i].pb, p.geo[i], s[i].d_img, s[i].d_atlas,
                                                s[i].d_atlas_label_norm, nullptr);
            const int rc = launch_atlas_pairs(
                &s[i].pb, s[i].d_atlas, s[i].d_atlas_label_norm, s[i].pb.n_atlas, t.scale,
                reinterpret_cast<f32x2*>(base + p.ws_off[i] +
                                         tiled_le_offset(&s[i].pb, p.geo[i].n_groups)),
                stream);
            if (rc) return rc;
        }
    }
    if (concurrent && hipEventRecord(events[0], stream) != hipSuccess) return PCT_ERR_HIP;
    for (int c = 0; c < nc; ++c) {
        const int hw = p.class_hw[c];
        const int sid = p.choice.shape_id[hw];
        const BatchKernel* bk = find_batch_kernel(sid);
        if (!bk) return PCT_ERR_UNSUPPORTED;
        hipStream_t st = (concurrent && c > 0) ? side[c - 1] : stream;
        if (concurrent && c > 0 && hipStreamWaitEvent(st, events[0], 0) != hipSuccess)
            return PCT_ERR_HIP;
        BatchArgs a;
        std::memset(&a, 0, sizeof(a));
        a.counters = counters + c * kChunks;
        int k = 0;
        for (int i : p.classes[c]) {
            a.sub[k] = make_tiled_args(&s[i].pb, p.geo[i], s[i].d_img, s[i].d_atlas,
                                       s[i].d_atlas_label_norm,
                                       reinterpret_cast<float*>(base + p.ws_off[i]));
            if (is_xterm_shape(sid))
                a.sub[k].le = reinterpret_cast<const f32x2*>(
                    base + p.ws_off[i] + tiled_le_offset(&s[i].pb, p.geo[i].n_groups));
            a.n_items[k] = p.geo[i].n_items;
            ++k;
        }
        a.n_sub = k;
        BatchedFn fn = bk->fn[hw - 1];
        void* kargs[] = {&a};
        if (hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3(grids[c]),
                            dim3(shapes()[sid].waves * 64), kargs, (size_t)p.class_lds[c],
                            st) != hipSuccess)
            return PCT_ERR_HIP;
        if (concurrent && c > 0) {
            if (hipEventRecord(events[c], st) != hipSuccess) return PCT_ERR_HIP;
            if (hipStreamWaitEvent(stream, events[c], 0) != hipSuccess) return PCT_ERR_HIP;
        }
    }
    for (int i = 0; i < p.n; ++i) {
        int rc;
        if (tiled_window(s[i].pb))
            rc = launch_finalize(&s[i].pb, p.geo[i],
                                 reinterpret_cast<const float*>(base + p.ws_off[i]), s[i].d_out,
                                 stream);
        else
            rc = launch_direct(&s[i].pb, s[i].d_img, s[i].d_atlas, s[i].d_atlas_label_norm,
                               s[i].d_out, stream);
        if (rc) return rc;
    }
    return PCT_OK;
}

// ---- configuration choice and graph cache -----------------------------------

struct BatchKey {
    std::vector<char> bytes;   // subject descriptors, workspace, configuration
    bool operator<(const BatchKey& o) const { return bytes < o.bytes; }
};

struct CachedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static std::mutex g_batch_mutex;
static std::map<std::vector<char>, pct_batch_choice> g_batch_tuned;   // shapes -> choice
static std::map<BatchKey, CachedGraph> g_graphs;

static std::vector<char> shape_signature(const pct_subject* s, int n) {
    std::vector<char> v;
    for (int i = 0; i < n; ++i) {
        const char* b = reinterpret_cast<const char*>(&s[i].pb);
        v.insert(v.end(), b, b + sizeof(pct_problem));
    }
    return v;
}

static pct_batch_choice empty_choice() {
    pct_batch_choice c;
    for (int h = 0; h < PCT_BATCH_CLASSES; ++h) {
        c.shape_id[h] = -1;
        c.atlas_group[h] = 1;
    }
    c.serial = 0;
    return c;
}

// Candidate atlas groups of a class: 1, 2 and all atlases of its largest
// subject.
static std::vector<int> class_groups(const pct_subject* s, int n, int hw) {
    int max_n = 1;
    for (int i = 0; i < n; ++i)
        if (has_items(s[i].pb) && class_of(s[i].pb) == hw)
            max_n = std::max(max_n, s[i].pb.n_atlas);
    std::vector<int> g = {1, 2, max_n};
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
    return g;
}

// Cost model choice per class: the shape and group with the least pooled item
// time over the subjects of the class (a persistent grid balances items
// across subjects, so the class costs its total item time over the slots).
// The classes run one after another: on MI355X concurrent class kernels
// measured slower (six subjects of 0.4-21 ms: 56.0 ms concurrent, 51.4 ms
// serial).
static pct_batch_choice model_choice(const pct_subject* s, int n) {
    pct_batch_choice ch = empty_choice();
    ch.serial = 1;
    for (int hw = 1; hw <= kMaxTiledHalfWindow; ++hw) {
        // the measured preference of the window where it has a persistent
        // kernel, else every batched shape
        const int pref = preferred_shape(hw);
        double best = -1;
        for (int pass = (pref >= 0 && find_batch_kernel(pref)) ? 0 : 1; pass < 2 && best < 0;
             ++pass)
            for (const auto& bk : batch_kernels()) {
                if (pass == 0 && bk.shape_id != pref) continue;
                for (int grp : class_groups(s, n, hw)) {
                    double cost = 0;
                    bool any = false, ok = true;
                    for (int i = 0; i < n && ok; ++i) {
                        if (!has_items(s[i].pb) || class_of(s[i].pb) != hw) continue;
                        TileGeometry g;
                        const double t = model_item_time(&s[i].pb, bk.shape_id,
                                                         std::min(grp, s[i].pb.n_atlas), &g);
                        if (t < 0) {
                            ok = false;
                            break;
                        }
                        cost += t * (double)g.n_items / ((double)kNumCu * g.wg_per_cu);
                        any = true;
                    }
                    if (!ok || !any) continue;
                    if (best < 0 || cost < best * 0.999) {
                        best = cost;
                        ch.shape_id[hw] = bk.shape_id;
                        ch.atlas_group[hw] = grp;
                    }
                }
            }
    }
    return ch;
}

static pct_batch_choice choose_batch(const pct_subject* s, int n) {
    {
        std::lock_guard<std::mutex> lk(g_batch_mutex);
        auto it = g_batch_tuned.find(shape_signature(s, n));
        if (it != g_batch_tuned.end()) return it->second;
    }
    return model_choice(s, n);
}

static void destroy_graph(CachedGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

static int run_batch(const pct_subject* s, int n, const pct_batch_choice& ch, void* ws,
                     size_t ws_bytes, bool own_ws, hipStream_t stream) {
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    BatchPlan p;
    if (!plan_batch(s, n, ch, &p)) return PCT_ERR_UNSUPPORTED;
    if (own_ws) {
        rc = ensure_workspace(dev, p.ws_bytes, &ws);
        if (rc) return rc;
    } else if (!ws || ws_bytes < p.ws_bytes || (reinterpret_cast<uintptr_t>(ws) & 255)) {
        return PCT_ERR_INVALID;
    }
    BatchKey key;
    {
        const char* b = reinterpret_cast<const char*>(s);
        key.bytes.assign(b, b + sizeof(pct_subject) * n);
        const char* c = reinterpret_cast<const char*>(&ch);
        key.bytes.insert(key.bytes.end(), c, c + sizeof(ch));
        const long long extra[2] = {(long long)reinterpret_cast<uintptr_t>(ws), dev};
        const char* e = reinterpret_cast<const char*>(extra);
        key.bytes.insert(key.bytes.end(), e, e + sizeof(extra));
    }
    std::lock_guard<std::mutex> lk(g_batch_mutex);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
        // capture on a private stream (the legacy default stream cannot be
        // captured), replay on the caller's stream
        const int nc = (int)p.classes.size();
        std::vector<int> grids(nc, 0);
        for (int c = 0; c < nc; ++c) {
            const int sid = p.choice.shape_id[p.class_hw[c]];
            const BatchKernel* bk = find_batch_kernel(sid);
            if (!bk) return PCT_ERR_UNSUPPORTED;
            grids[c] = persistent_grid(dev, sid, bk->fn[p.class_hw[c] - 1], p.class_lds[c]);
            if (grids[c] < 1) return PCT_ERR_HIP;
        }
        hipStream_t cap = nullptr;
        if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) return PCT_ERR_HIP;
        std::vector<hipStream_t> side(nc > 1 ? nc - 1 : 0, nullptr);
        std::vector<hipEvent_t> events(nc + 1, nullptr);
        rc = PCT_OK;
        for (auto& st : side)
            if (!rc && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
                rc = PCT_ERR_HIP;
        for (auto& ev : events)
            if (!rc && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
                rc = PCT_ERR_HIP;
        CachedGraph cg;
        if (!rc && hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = PCT_ERR_HIP;
        } else if (!rc) {
            rc = enqueue_batch(s, p, grids, ws, cap, side, events);
            if (hipStreamEndCapture(cap, &cg.graph) != hipSuccess && !rc) rc = PCT_ERR_HIP;
        }
        if (!rc && hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0) != hipSuccess)
            rc = PCT_ERR_HIP;
        for (auto& ev : events)
            if (ev) (void)hipEventDestroy(ev);
        for (auto& st : side)
            if (st) (void)hipStreamDestroy(st);
        (void)hipStreamDestroy(cap);
        if (rc) {
            (void)hipGetLastError();
            destroy_graph(cg);
            return rc;
        }
        if (g_graphs.size() >= 32) {   // bounded cache
            for (auto& kv : g_graphs) destroy_graph(kv.second);
            g_graphs.clear();
        }
        it = g_graphs.emplace(key, cg).first;
    }
    if (hipGraphLaunch(it->second.exec, stream) != hipSuccess) {
        (void)hipGetLastError();
        return PCT_ERR_HIP;
    }
    return PCT_OK;
}

// Mean time of `reps` runs after one warm-up run (which captures the graph).
static int time_batch(const pct_subject* s, int n, const pct_batch_choice& ch, int reps,
                      hipStream_t st, hipEvent_t e0, hipEvent_t e1, double* ms) {
    int rc = run_batch(s, n, ch, nullptr, 0, true, st);
    if (rc) return rc;
    if (hipEventRecord(e0, st) != hipSuccess) return PCT_ERR_HIP;
    for (int i = 0; i < reps && !rc; ++i) rc = run_batch(s, n, ch, nullptr, 0, true, st);
    if (hipEventRecord(e1, st) != hipSuccess) return PCT_ERR_HIP;
    if (hipEventSynchronize(e1) != hipSuccess) return PCT_ERR_HIP;
    if (rc) return rc;
    float t = 0;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return PCT_ERR_HIP;
    *ms = t / reps;
    return PCT_OK;
}

}  // namespace pct

using namespace pct;

extern "C" size_t pct_fuse_batch_workspace_size(const pct_subject* subjects, int n_subjects) {
    if (validate_batch(subjects, n_subjects)) return 0;
    BatchPlan p;
    if (!plan_batch(subjects, n_subjects, choose_batch(subjects, n_subjects), &p)) return 0;
    return p.ws_bytes;
}

extern "C" int pct_fuse_batch_dev(const pct_subject* subjects, int n_subjects,
                                  pct_stream_t stream) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    return run_batch(subjects, n_subjects, choose_batch(subjects, n_subjects), nullptr, 0, true,
                     reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_fuse_batch_dev_ws(const pct_subject* subjects, int n_subjects, void* d_ws,
                                     size_t ws_bytes, pct_stream_t stream) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    return run_batch(subjects, n_subjects, choose_batch(subjects, n_subjects), d_ws, ws_bytes,
                     false, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_fuse_batch_choice_get(const pct_subject* subjects, int n_subjects,
                                         pct_batch_choice* choice) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    if (!choice) return PCT_ERR_INVALID;
    *choice = choose_batch(subjects, n_subjects);
    return PCT_OK;
}

extern "C" int pct_fuse_batch_choice_set(const pct_subject* subjects, int n_subjects,
                                         const pct_batch_choice* choice) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    if (!choice || (choice->serial != 0 && choice->serial != 1)) return PCT_ERR_INVALID;
    pct_batch_choice ch = empty_choice();
    ch.serial = choice->serial;
    for (int i = 0; i < n_subjects; ++i) {
        if (!has_items(subjects[i].pb)) continue;
        const int hw = class_of(subjects[i].pb);
        if (!find_batch_kernel(choice->shape_id[hw]) || choice->atlas_group[hw] < 1)
            return PCT_ERR_INVALID;
        ch.shape_id[hw] = choice->shape_id[hw];
        ch.atlas_group[hw] = choice->atlas_group[hw];
    }
    BatchPlan p;
    if (!plan_batch(subjects, n_subjects, ch, &p)) return PCT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_batch_mutex);
    g_batch_tuned[shape_signature(subjects, n_subjects)] = ch;
    return PCT_OK;
}

extern "C" int pct_fuse_batch_config(const pct_subject* subjects, int n_subjects, int shape_id,
                                     int atlas_group) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    if (!find_batch_kernel(shape_id) || atlas_group < 1) return PCT_ERR_INVALID;
    pct_batch_choice ch = empty_choice();
    for (int h = 1; h <= kMaxTiledHalfWindow; ++h) {
        ch.shape_id[h] = shape_id;
        ch.atlas_group[h] = atlas_group;
    }
    return pct_fuse_batch_choice_set(subjects, n_subjects, &ch);
}

extern "C" int pct_autotune_batch(const pct_subject* subjects, int n_subjects, int reps,
                                  pct_stream_t stream, pct_tile_config* best) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (reps < 1) reps = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return PCT_ERR_HIP;
    }
    pct_batch_choice ch = empty_choice();
    double class_ms[PCT_BATCH_CLASSES] = {0};
    // every class alone: its subjects, every shape and group
    for (int hw = 1; hw <= kMaxTiledHalfWindow && !rc; ++hw) {
        std::vector<pct_subject> cls;
        for (int i = 0; i < n_subjects; ++i)
            if (has_items(subjects[i].pb) && class_of(subjects[i].pb) == hw)
                cls.push_back(subjects[i]);
        if (cls.empty()) continue;
        double best_ms = -1;
        for (const auto& bk : batch_kernels()) {
            for (int grp : class_groups(cls.data(), (int)cls.size(), hw)) {
                pct_batch_choice c = empty_choice();
                c.shape_id[hw] = bk.shape_id;
                c.atlas_group[hw] = grp;
                BatchPlan p;
                if (!plan_batch(cls.data(), (int)cls.size(), c, &p)) continue;
                double ms = 0;
                rc = time_batch(cls.data(), (int)cls.size(), c, reps, st, e0, e1, &ms);
                if (rc) break;
                if (best_ms < 0 || ms < best_ms) {
                    best_ms = ms;
                    ch.shape_id[hw] = bk.shape_id;
                    ch.atlas_group[hw] = grp;
                }
            }
            if (rc) break;
        }
        if (!rc && best_ms < 0) rc = PCT_ERR_UNSUPPORTED;
        class_ms[hw] = best_ms;
    }
    // the winners together: classes concurrent, then one after another
    double batch_ms = -1;
    for (int serial = 0; serial <= 1 && !rc; ++serial) {
        pct_batch_choice c = ch;
        c.serial = serial;
        double ms = 0;
        rc = time_batch(subjects, n_subjects, c, reps, st, e0, e1, &ms);
        if (!rc && (batch_ms < 0 || ms < batch_ms)) {
            batch_ms = ms;
            ch.serial = serial;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_batch_mutex);
        g_batch_tuned[shape_signature(subjects, n_subjects)] = ch;
    }
    if (best) {
        int top = -1;
        for (int hw = 1; hw <= kMaxTiledHalfWindow; ++hw)
            if (ch.shape_id[hw] >= 0 && (top < 0 || class_ms[hw] > class_ms[top])) top = hw;
        std::memset(best, 0, sizeof(*best));
        BatchPlan p;
        plan_batch(subjects, n_subjects, ch, &p);
        for (int l : p.class_lds) best->lds_bytes = std::max(best->lds_bytes, l);
        for (int i = 0; i < n_subjects; ++i) best->items += p.geo[i].n_items;
        best->id = top > 0 ? ch.shape_id[top] : -1;
        if (top > 0) {
            const TileShape& sh = shapes()[ch.shape_id[top]];
            best->seg_width = sh.seg_width;
            best->bx = sh.bx;
            best->bz = sh.bz;
            best->waves = sh.waves;
            best->form = sh.form;
            best->atlas_group = ch.atlas_group[top];
        }
        best->est_cost = -1;
        best->measured_ms = batch_ms;
    }
    return PCT_OK;
}

namespace pct {
void release_batch_graphs() {
    std::lock_guard<std::mutex> lk(g_batch_mutex);
    for (auto& kv : g_graphs) destroy_graph(kv.second);
    g_graphs.clear();
}
}  // namespace pct
