// Batched fusion: many problems (subjects) of uneven volume size, atlas
// count and search window in one persistent kernel.
//
// The work items of the subjects - (tile, atlas group) pairs of the tiled
// kernel - form one queue per search-window class (half window HW), subjects
// with the longest items first.  For every class present, a persistent grid
// of as many workgroups as the GPU holds at once pulls items from its queue
// with a device-scope atomic counter until the queue is empty; each item runs
// the same tiled_item body as the single-problem kernel, fully specialised
// on HW (one kernel per class keeps the register allocation of that window
// instead of the union over all windows).  A subject's result is therefore
// bit-identical to its individual run with the same tile shape and atlas
// group.  The class kernels are concurrent branches of one captured hipGraph,
// the longest class first, so the tail of one queue overlaps the next; the
// counter reset, the direct kernel for windows without a tiled
// specialisation and the per-subject finalize passes complete the graph,
// which is replayed while the subject descriptors stay the same.
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

struct BatchArgs {
    TiledArgs sub[PCT_MAX_BATCH];            // subjects of one window class, in queue order
    long long begin[PCT_MAX_BATCH + 1];      // first queue position of each subject
    int n_sub;
    unsigned long long* counter;
};
static_assert(sizeof(BatchArgs) <= 4096, "batch descriptor must fit the kernel arguments");

template <int W, int BX, int BZ, int NW, int HW>
__global__ __launch_bounds__(NW * 64) void fuse_batched(BatchArgs args) {
    extern __shared__ float lds[];
    __shared__ long long next_item;
    const long long total = args.begin[args.n_sub];
    for (;;) {
        if (threadIdx.x == 0) next_item = (long long)atomicAdd(args.counter, 1ULL);
        __syncthreads();
        const long long it = next_item;
        __syncthreads();
        if (it >= total) break;   // every workgroup leaves once the queue is drained
        int k = 0;
        while (k + 1 < args.n_sub && args.begin[k + 1] <= it) ++k;
        tiled_item<W, BX, BZ, NW, HW>(args.sub[k], (int)(it - args.begin[k]), lds);
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

// Tile shapes with persistent instantiations (ids of shapes()).
static const std::vector<BatchKernel>& batch_kernels() {
    static const std::vector<BatchKernel> k = {
        make_batch_kernel<32, 4, 2, 4>(8),
        make_batch_kernel<32, 2, 2, 8>(10),
        make_batch_kernel<32, 2, 2, 4>(11),
        make_batch_kernel<32, 2, 4, 8>(13),
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

// Layout of a batch in the workspace and in the queues for one configuration.
struct BatchPlan {
    int shape_id = -1, group = 1;
    int n = 0;
    std::vector<TileGeometry> geo;            // per subject (input order)
    std::vector<size_t> ws_off;               // per subject, bytes from the workspace start
    std::vector<std::vector<int>> classes;    // tiled subjects by window class, longest first
    std::vector<int> class_hw;
    std::vector<int> class_lds;
    size_t ws_bytes = 0;
};

static bool plan_batch(const pct_subject* s, int n, int shape_id, int group, BatchPlan* p) {
    p->shape_id = shape_id;
    p->group = group;
    p->n = n;
    p->geo.assign(n, TileGeometry());
    p->ws_off.assign(n, 0);
    p->classes.clear();
    p->class_hw.clear();
    p->class_lds.clear();
    size_t off = 256;   // the queue counters, one per window class
    std::vector<double> item_cost(n, 0.0);
    std::vector<int> by_hw[kMaxTiledHalfWindow + 1];
    for (int i = 0; i < n; ++i) {
        if (!tiled_window(s[i].pb)) continue;
        TileGeometry g;
        if (!make_geometry(&s[i].pb, shape_id, std::min(group, s[i].pb.n_atlas), &g))
            return false;
        p->geo[i] = g;
        p->ws_off[i] = off;
        off += tiled_ws_bytes(&s[i].pb, g.n_groups);
        item_cost[i] = g.group;
        if (g.n_items > 0) by_hw[g.hw].push_back(i);
    }
    // classes with the longest items first; inside a class, longest items first
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

// Enqueues one batched run on `stream`, which is being captured: the class
// kernels fork onto `side` streams and join back before the finalize passes.
static int enqueue_batch(const pct_subject* s, const BatchPlan& p, const std::vector<int>& grids,
                         void* ws, hipStream_t stream, std::vector<hipStream_t>& side,
                         std::vector<hipEvent_t>& events) {
    const BatchKernel* bk = find_batch_kernel(p.shape_id);
    if (!bk) return PCT_ERR_UNSUPPORTED;
    char* base = static_cast<char*>(ws);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(base);
    const int nc = (int)p.classes.size();
    hipLaunchKernelGGL(reset_queues, dim3(1), dim3(64), 0, stream, counters, nc);
    if (hipGetLastError() != hipSuccess) return PCT_ERR_HIP;
    if (hipEventRecord(events[0], stream) != hipSuccess) return PCT_ERR_HIP;
    for (int c = 0; c < nc; ++c) {
        hipStream_t st = c == 0 ? stream : side[c - 1];
        if (c > 0 && hipStreamWaitEvent(st, events[0], 0) != hipSuccess) return PCT_ERR_HIP;
        BatchArgs a;
        std::memset(&a, 0, sizeof(a));
        a.counter = counters + c;
        long long pos = 0;
        int k = 0;
        for (int i : p.classes[c]) {
            a.sub[k] = make_tiled_args(&s[i].pb, p.geo[i], s[i].d_img, s[i].d_atlas,
                                       s[i].d_atlas_label_norm,
                                       reinterpret_cast<float*>(base + p.ws_off[i]));
            a.begin[k] = pos;
            pos += p.geo[i].n_items;
            ++k;
        }
        a.begin[k] = pos;
        a.n_sub = k;
        BatchedFn fn = bk->fn[p.class_hw[c] - 1];
        void* kargs[] = {&a};
        if (hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3(grids[c]),
                            dim3(shapes()[p.shape_id].waves * 64), kargs,
                            (size_t)p.class_lds[c], st) != hipSuccess)
            return PCT_ERR_HIP;
        if (c > 0) {
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
static std::map<std::vector<char>, std::pair<int, int>> g_batch_tuned;   // shapes -> config
static std::map<BatchKey, CachedGraph> g_graphs;

static std::vector<char> shape_signature(const pct_subject* s, int n) {
    std::vector<char> v;
    for (int i = 0; i < n; ++i) {
        const char* b = reinterpret_cast<const char*>(&s[i].pb);
        v.insert(v.end(), b, b + sizeof(pct_problem));
    }
    return v;
}

// Model choice: the shape and group with the least summed single-problem
// cost estimate (the batch shares one shape and one group).
static bool choose_batch(const pct_subject* s, int n, int* shape_id, int* group) {
    {
        std::lock_guard<std::mutex> lk(g_batch_mutex);
        auto it = g_batch_tuned.find(shape_signature(s, n));
        if (it != g_batch_tuned.end()) {
            *shape_id = it->second.first;
            *group = it->second.second;
            return true;
        }
    }
    *shape_id = batch_kernels().front().shape_id;
    *group = 1;
    return true;
}

static void destroy_graph(CachedGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

static int run_batch(const pct_subject* s, int n, int shape_id, int group, void* ws,
                     size_t ws_bytes, bool own_ws, hipStream_t stream) {
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    BatchPlan p;
    if (!plan_batch(s, n, shape_id, group, &p)) return PCT_ERR_UNSUPPORTED;
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
        const long long extra[4] = {shape_id, group, (long long)reinterpret_cast<uintptr_t>(ws),
                                    dev};
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
        const BatchKernel* bk = find_batch_kernel(p.shape_id);
        if (!bk) return PCT_ERR_UNSUPPORTED;
        for (int c = 0; c < nc; ++c) {
            grids[c] = persistent_grid(dev, p.shape_id, bk->fn[p.class_hw[c] - 1], p.class_lds[c]);
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

}  // namespace pct

using namespace pct;

extern "C" size_t pct_fuse_batch_workspace_size(const pct_subject* subjects, int n_subjects) {
    if (validate_batch(subjects, n_subjects)) return 0;
    int sid = 0, grp = 1;
    choose_batch(subjects, n_subjects, &sid, &grp);
    BatchPlan p;
    if (!plan_batch(subjects, n_subjects, sid, grp, &p)) return 0;
    return p.ws_bytes;
}

extern "C" int pct_fuse_batch_dev(const pct_subject* subjects, int n_subjects,
                                  pct_stream_t stream) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    int sid = 0, grp = 1;
    choose_batch(subjects, n_subjects, &sid, &grp);
    return run_batch(subjects, n_subjects, sid, grp, nullptr, 0, true,
                     reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_fuse_batch_dev_ws(const pct_subject* subjects, int n_subjects, void* d_ws,
                                     size_t ws_bytes, pct_stream_t stream) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    int sid = 0, grp = 1;
    choose_batch(subjects, n_subjects, &sid, &grp);
    return run_batch(subjects, n_subjects, sid, grp, d_ws, ws_bytes, false,
                     reinterpret_cast<hipStream_t>(stream));
}

extern "C" int pct_fuse_batch_config(const pct_subject* subjects, int n_subjects, int shape_id,
                                     int atlas_group) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    if (!find_batch_kernel(shape_id) || atlas_group < 1) return PCT_ERR_INVALID;
    BatchPlan p;
    if (!plan_batch(subjects, n_subjects, shape_id, atlas_group, &p)) return PCT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_batch_mutex);
    g_batch_tuned[shape_signature(subjects, n_subjects)] = {shape_id, atlas_group};
    return PCT_OK;
}

extern "C" int pct_autotune_batch(const pct_subject* subjects, int n_subjects, int reps,
                                  pct_stream_t stream, pct_tile_config* best) {
    int rc = validate_batch(subjects, n_subjects);
    if (rc) return rc;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (reps < 1) reps = 1;
    int max_n = 1;
    for (int i = 0; i < n_subjects; ++i) max_n = std::max(max_n, subjects[i].pb.n_atlas);
    std::vector<int> groups = {1, 2, max_n};
    std::sort(groups.begin(), groups.end());
    groups.erase(std::unique(groups.begin(), groups.end()), groups.end());
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return PCT_ERR_HIP;
    double best_ms = -1;
    int best_sid = -1, best_grp = 1;
    for (const auto& bk : batch_kernels()) {
        for (int grp : groups) {
            BatchPlan p;
            if (!plan_batch(subjects, n_subjects, bk.shape_id, grp, &p)) continue;
            rc = run_batch(subjects, n_subjects, bk.shape_id, grp, nullptr, 0, true, s);  // warm-up
            if (rc) break;
            (void)hipEventRecord(e0, s);
            for (int i = 0; i < reps && !rc; ++i)
                rc = run_batch(subjects, n_subjects, bk.shape_id, grp, nullptr, 0, true, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = PCT_ERR_HIP;
            if (rc) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= reps;
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best_sid = bk.shape_id;
                best_grp = grp;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_sid < 0) return PCT_ERR_UNSUPPORTED;
    {
        std::lock_guard<std::mutex> lk(g_batch_mutex);
        g_batch_tuned[shape_signature(subjects, n_subjects)] = {best_sid, best_grp};
    }
    if (best) {
        const TileShape& sh = shapes()[best_sid];
        std::memset(best, 0, sizeof(*best));
        best->id = best_sid;
        best->seg_width = sh.seg_width;
        best->bx = sh.bx;
        best->bz = sh.bz;
        best->waves = sh.waves;
        best->atlas_group = best_grp;
        BatchPlan p;
        plan_batch(subjects, n_subjects, best_sid, best_grp, &p);
        best->lds_bytes = 0;
        for (int l : p.class_lds) best->lds_bytes = std::max(best->lds_bytes, l);
        long long items = 0;
        for (int i = 0; i < n_subjects; ++i) items += p.geo[i].n_items;
        best->items = items;
        best->est_cost = -1;
        best->measured_ms = best_ms;
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
