/* pct.h - C ABI of the MI355X (gfx950) multi-atlas patch-fusion library.
 *
 * Non-local multi-atlas patch fusion for pseudo-CT synthesis: for every
 * interior target voxel v, every atlas a and every offset d of the NH^3
 * search window,
 *     SSD = sum_{p in [-1,1]^3} (T(v+p) - A_a(v+d+p))^2
 *     w   = exp(-SSD / div),  div = 2*27*beta*sigma^2
 *     out(v) = sum w*L'_a(v+d) / sum w
 * with L' the atlas CT normalised to [0,255] (see pct_normalize_labels).
 *
 * Layout: float32 volumes, linear index z*nx*ny + x*ny + y (y fastest);
 * atlas a starts at a*nx*ny*nz.  Identical to the reference program
 * (src/CPU/NewPatchedBrainPureC.cpp:1288, 1300; src/GPU/PatchBasedPseudoCT.cu:169).
 *
 * Every entry point returns PCT_OK (0) or a negative pct_status; the library
 * never calls exit().  Nothing here holds global problem state: dimensions
 * travel in pct_problem instead of the reference's globals nx, ny, nz, num,
 * nh (cpp:81-82, cu:652-654).
 */
#ifndef PCT_PCT_H
#define PCT_PCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCT_ABI_VERSION 2

/* Same object as hipStream_t; declared here so that the header needs no HIP
 * headers.  NULL means the default stream. */
typedef struct ihipStream_t* pct_stream_t;

typedef enum pct_status {
    PCT_OK = 0,
    PCT_ERR_INVALID = -1,     /* bad argument (null pointer, NH < 1, patch != 3, ...) */
    PCT_ERR_NO_DEVICE = -2,   /* no HIP device, or the device is not gfx950 */
    PCT_ERR_HIP = -3,         /* a HIP runtime call failed */
    PCT_ERR_NO_MEMORY = -4,   /* host or device allocation failed */
    PCT_ERR_IO = -5,          /* file could not be opened / written */
    PCT_ERR_FORMAT = -6,      /* malformed MatrixIO / configuration file */
    PCT_ERR_DIMS = -7,        /* dimensions of the input files do not match */
    PCT_ERR_UNSUPPORTED = -8, /* no kernel configuration fits the request */
    PCT_ERR_TOO_LARGE = -9    /* request exceeds a documented maximum */
} pct_status;

/* EXEC_CHOICE of the GPU program (src/GPU/PatchBasedPseudoCT.cu:71-78).
 * The reference's three kernels compute the same function and differ in
 * how they use the memory hierarchy (and in whether the patch difference is
 * rounded to float before widening, <= 1 float ulp per term).  Here GM, GM2
 * and SM all run the LDS-tiled kernel with separable box-sum reuse, in FP32,
 * for NH up to 15.  DIRECT has no counterpart in the reference (value 3 is
 * unhandled there, cu:2026-2205): the one-lane-per-voxel kernel, 27-term
 * SSD per candidate, an independent implementation used for cross-checks;
 * NH above 15 runs it for every choice.  GPU_TEST_TIME is a CLI mode. */
typedef enum pct_exec_choice {
    PCT_CHOICE_GM = 0,
    PCT_CHOICE_GM2 = 1,
    PCT_CHOICE_SM = 2,
    PCT_CHOICE_GPU_TEST_TIME = 4,
    PCT_CHOICE_DIRECT = 5
} pct_exec_choice;

typedef struct pct_problem {
    int nx, ny, nz;   /* volume dimensions (MatrixIO rows, cols, depth) */
    int n_atlas;      /* number of atlas pairs (reference global `num`) */
    int nh;           /* search window edge NH; half window = (NH-1)/2 (cpp:1243) */
    int patch;        /* patch edge; must be 3 (cpp:1240) */
    double sigma;     /* cpp:1560 */
    double beta;      /* cpp:1561 */
    double div;       /* > 0: the weight divisor itself, as the reference's
                         wrappers take it (cu:1635, cpp:1350), and sigma, beta
                         are not used; 0: div = pct_weight_divisor(sigma, beta) */
} pct_problem;

const char* pct_strerror(int status);
int pct_abi_version(void);

/* ---- pre-processing (src/CPU/NewPatchedBrainPureC.cpp:1583-1601) ---- */

/* div = 2*27*beta*sigma^2 in float64 (cpp:1601, cu:2154). */
double pct_weight_divisor(double sigma, double beta);

/* In place, float32: L' = 255.0f*((L+1024.0f)/4095.0f) (cpp:1583-1593,
 * cu:2009-2020). Host buffer. */
void pct_normalize_labels(float* label, size_t n);

/* Same operation on a device buffer, on `stream`. */
int pct_normalize_labels_dev(float* d_label, size_t n, pct_stream_t stream);

/* ---- the hot path ---- */

/* Replaces MultiPatchBasedSegmentationGpu (src/GPU/PatchBasedPseudoCT.cu:1635)
 * and its CPU twin MultiAtlasPatchBasedSegmentationOpenMP (cpp:1350).
 * Host buffers; copies in, runs and copies `out` back like the reference
 * wrapper (cu:1687-1744), and returns when `out` is written.  The atlas
 * copies go in chunks that overlap the fusion of the chunks already copied.
 * Copies in both directions pass through 64 MiB of pinned host memory
 * (four 16 MiB buffers, filled and emptied by the OpenMP threads).  The
 * device buffers and the pinned buffers stay allocated for the next
 * host-buffer call (grown on demand; pct_release frees them); host-buffer
 * calls on one device run one at a time.  `out` receives every voxel:
 * interior = fused value (NaN where |P| or |S| <= DBL_EPSILON,
 * cpp:1437-1443), border = 0 (cu:1693).  `choice` is a pct_exec_choice
 * (GM, GM2, SM or DIRECT).  The output equals pct_fuse_dev's bit for bit. */
int pct_fuse_host(const pct_problem* pb, const float* img, const float* atlas,
                  const float* atlas_label_norm, float* out, int choice);

/* The whole SM/GM/GM2 block of the reference's GPU main (cu:2152-2179): the
 * fusion of pct_fuse_host, then RegularizationStep, RoundMatrix and the HU
 * rescale (cpp:1059-1106, 1221-1227, cu:2166-2179) applied on the device
 * before the single copy back.  `out_hu` equals pct_fuse_host followed by
 * pct_postprocess bit for bit.  Volumes of nx*ny*nz >= 2^31 - 1 voxels
 * return PCT_ERR_TOO_LARGE (the device post chain's limit). */
int pct_pseudo_ct_host(const pct_problem* pb, const float* img, const float* atlas,
                       const float* atlas_label_norm, float* out_hu, int choice);

/* Device-pointer variant: inputs already resident in HBM, enqueued on
 * `stream`, asynchronous.  Uses a per-device workspace owned by the library
 * (grown on demand; calls that overlap on one device must be ordered on one
 * stream).  No host synchronisation unless the workspace has to grow. */
int pct_fuse_dev(const pct_problem* pb, const float* d_img, const float* d_atlas,
                 const float* d_atlas_label_norm, float* d_out, int choice,
                 pct_stream_t stream);

/* Bytes of device workspace pct_fuse_dev_ws needs for this problem/choice. */
size_t pct_fuse_workspace_size(const pct_problem* pb, int choice);

/* Same as pct_fuse_dev with a caller-provided workspace of at least
 * pct_fuse_workspace_size() bytes (256-byte aligned). */
int pct_fuse_dev_ws(const pct_problem* pb, const float* d_img, const float* d_atlas,
                    const float* d_atlas_label_norm, float* d_out, int choice,
                    void* d_workspace, size_t workspace_bytes, pct_stream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pct_fuse_dev with no allocation of any kind.  Its partial
 * sums live in a 64 MiB array that the library's code object holds on every
 * GPU it runs on until the process ends; it runs one atlas group with the
 * squared-difference kernels, so the tiled choices need 8 * nx*ny*nz bytes
 * of it (at most 8 388 608 voxels, e.g. 256x256x128; larger volumes return
 * PCT_ERR_TOO_LARGE) and DIRECT needs none.  Output as pct_fuse_dev's within
 * the FP32 tolerance (another tile shape and atlas grouping).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and must be ordered by the caller. */
int pct_fuse_dev_noalloc(const pct_problem* pb, const float* d_img, const float* d_atlas,
                         const float* d_atlas_label_norm, float* d_out, int choice,
                         pct_stream_t stream);

/* ---- kernel configuration / autotuner ---- */

/* Description of the tiled (SM) kernel configuration chosen for a problem. */
typedef struct pct_tile_config {
    int id;            /* index into the library's configuration table */
    int seg_width;     /* lanes per y segment (outputs per segment = seg_width-2) */
    int bx, bz;        /* per-lane output block in x and z */
    int waves;         /* wavefronts per workgroup */
    int atlas_group;   /* atlases per work item */
    int lds_bytes;     /* dynamic LDS per workgroup */
    long long items;   /* work items (workgroups) launched */
    double est_cost;   /* cost-model estimate (arbitrary units) */
    double measured_ms;/* <0 if not measured */
    int form;          /* 0: squared differences per candidate; 1: cross terms
                          (2 T'A' products, per-voxel 2^-AA folded into the
                          staged CT; squared differences where the value range
                          needs them) */
} pct_tile_config;

/* Number of tiled-kernel configurations compiled into the library. */
int pct_tile_config_count(void);

/* Cost-model choice (no GPU work). */
int pct_tile_config_choose(const pct_problem* pb, pct_tile_config* out);

/* Measured choice: runs every admissible configuration `reps` times on the
 * given device inputs, keeps the fastest, and caches the decision for this
 * problem shape for the rest of the process (later pct_fuse_* calls with the
 * same shape use it).  Synchronises `stream`. */
int pct_autotune(const pct_problem* pb, const float* d_img, const float* d_atlas,
                 const float* d_atlas_label_norm, float* d_out, int reps,
                 pct_stream_t stream, pct_tile_config* best);

/* Force a configuration (by id, atlas_group) for a problem shape; -1 clears. */
int pct_tile_config_force(int id, int atlas_group);

/* ---- batched fusion (no counterpart in the reference, whose wrapper fuses
 * one subject per call, cu:1635; for callers with many subjects) ---- */

#define PCT_MAX_BATCH 16

/* One subject of a batch: its problem and its device buffers. */
typedef struct pct_subject {
    pct_problem pb;
    const float* d_img;
    const float* d_atlas;
    const float* d_atlas_label_norm;
    float* d_out;
} pct_subject;

/* Fuses up to PCT_MAX_BATCH subjects of any sizes, atlas counts and windows
 * in one persistent kernel fed by a work queue (plus the finalize passes),
 * captured once into a hipGraph and replayed while the descriptors stay the
 * same.  Each subject's output equals its individual pct_fuse_dev(SM) run
 * with the same tile configuration bit for bit.  Asynchronous on `stream`;
 * uses the library's per-device workspace. */
int pct_fuse_batch_dev(const pct_subject* subjects, int n_subjects, pct_stream_t stream);

/* Bytes of workspace pct_fuse_batch_dev_ws needs. */
size_t pct_fuse_batch_workspace_size(const pct_subject* subjects, int n_subjects);

/* pct_fuse_batch_dev with a caller-provided workspace (256-byte aligned). */
int pct_fuse_batch_dev_ws(const pct_subject* subjects, int n_subjects, void* d_workspace,
                          size_t workspace_bytes, pct_stream_t stream);

/* Measures, for every window class of these subjects, the batched tile
 * shapes and atlas groups on that class alone, then the winners with the
 * classes concurrent and one after another; keeps the fastest for this set
 * of problem shapes.  `best` reports the configuration of the class that
 * took longest and the measured time of the whole batch.  Synchronises. */
int pct_autotune_batch(const pct_subject* subjects, int n_subjects, int reps,
                       pct_stream_t stream, pct_tile_config* best);

/* Sets one tile shape (pct_tile_config.id) and atlas group for every window
 * class of this set of problem shapes, the classes running concurrently. */
int pct_fuse_batch_config(const pct_subject* subjects, int n_subjects, int shape_id,
                          int atlas_group);

/* The configuration of a batch per window class, indexed by the half window
 * (NH-1)/2 = 1..7 (index 0 unused). */
#define PCT_BATCH_CLASSES 8
typedef struct pct_batch_choice {
    int shape_id[PCT_BATCH_CLASSES];     /* -1: no tiled subject of that class */
    int atlas_group[PCT_BATCH_CLASSES];
    int serial;                          /* 1: class kernels one after another; 0: concurrent */
} pct_batch_choice;

/* The configuration pct_fuse_batch_dev uses for these subjects (tuned, set
 * or the cost model's). */
int pct_fuse_batch_choice_get(const pct_subject* subjects, int n_subjects,
                              pct_batch_choice* choice);

/* Sets the configuration of every window class present in `subjects`. */
int pct_fuse_batch_choice_set(const pct_subject* subjects, int n_subjects,
                              const pct_batch_choice* choice);

/* ---- learned candidate proposal (no counterpart in the reference, which
 * always searches all n*NH^3 candidates, cpp:1399-1432; for callers that
 * trade a bounded HU error for speed) ----
 *
 * A per-voxel MLP 30 -> 64 -> 64 -> 16 (ReLU after the first two layers)
 * embeds target and atlas voxels; candidates whose embeddings are close to
 * the target's are kept.  Parameter vector, float32, PCT_MLP_NPARAMS values:
 *   W1[32][64] (input-major; input rows 30, 31 unused), b1[64],
 *   W2[64][64], b2[64], W3[64][16], b3[16], mu, inv_sd.
 * Features of voxel u of volume V, with g(u) = (V(u) - mu) * inv_sd inside
 * the volume and 0 outside:
 *   f[9*pz+3*px+py] = g(u + (pz-1, px-1, py-1)), then
 *   f[27..29] = (g(u+e) - g(u-e)) / 2 for e = z, x, y.
 * Inference runs in BF16 (weights and activations) with FP32 accumulation
 * on MFMA; training keeps FP32 master weights. */
#define PCT_MLP_NPARAMS 7314

/* He-initialised parameters (host), seeded; mu / inv_sd as given. */
int pct_mlp_init(float* params, uint64_t seed, float mu, float inv_sd);

/* 16-dim BF16 embedding (raw 16-bit words, [voxel][16]) of every voxel of
 * a device volume; `params` is a host parameter vector. */
int pct_embed_dev(const float* params, const float* d_vol, int nx, int ny, int nz,
                  uint16_t* d_emb, pct_stream_t stream);

/* Options of the pruned fusion. */
typedef struct pct_pruned_opts {
    int topk;        /* candidate offsets kept per tile and atlas (32) */
    int samples;     /* offsets on a fixed lattice weighted exactly to estimate
                        the mean weight of the candidates that are not kept */
    float margin;    /* tiles whose gap between the topk-th and the next
                        embedding score is below margin times the score
                        spread (median - best) take the full search; 0 = none,
                        negative = every tile (validation) */
    int shape_id;    /* tile shape (pct_tile_config.id); -1 = default */
    int max_slabs;   /* kept offsets come from at most this many dz planes of
                        the window: the lattice's and the best-scoring others
                        (bounds the LDS staging per tile); 0 = any */
    int embed_all;   /* 1: embed every atlas voxel instead of only the probe
                        windows the selection reads (same result; checking) */
    float max_err;   /* > 0: tiles in which some voxel's estimated error
                        exceeds max_err (label units; 1 = 16.06 HU) take the
                        full search.  The tail term replaces the weighted CT
                        mean of the candidates it stands for by their plain
                        mean, which misses by cov(w, L) / mean(w) over them;
                        the estimate is that ratio over the lattice samples'
                        exact weights, times the tail's share of the weight
                        sum.  <= 0: no estimate */
} pct_pruned_opts;

/* Learned-pruned fusion.  Per tile of target voxels and per atlas, the
 * topk candidate offsets with the smallest embedding distance (summed over a
 * compact probe cluster at the tile centre) are weighted exactly in FP32 for
 * every voxel of the tile; the remaining candidates enter through their CT
 * window sum times the mean exact weight of `samples` lattice offsets.  Tiles
 * with a low selection margin, or (max_err > 0) with a voxel whose estimated
 * error exceeds max_err, run the full exact search and give exactly the
 * plain kernel's sums for that tile and atlas.  Output as
 * pct_fuse_dev.  If fallback_frac is not NULL the call synchronises and
 * stores the fraction of (tile, atlas) items that fell back.  Planes of
 * nx*ny >= 2^31 voxels return PCT_ERR_TOO_LARGE.  Part of the work runs on
 * a second stream of the library's, forked from and joined back to
 * `stream` with events: all of it comes after the work already queued on
 * `stream` and before the work queued after the call.  The call copies
 * small tables from host memory, so it is not meant for graph capture. */
int pct_fuse_pruned_dev(const pct_problem* pb, const float* params, const pct_pruned_opts* opts,
                        const float* d_img, const float* d_atlas,
                        const float* d_atlas_label_norm, float* d_out, double* fallback_frac,
                        pct_stream_t stream);

/* Training objective.  An anchor is an interior target voxel u and an atlas
 * a.  Over the NH^3 window offsets c, the teacher distribution is
 * p_c = softmax_c(-SSD_c / T), SSD_c the SSD of the 3x3x3 MRI patches of the
 * target at u and of atlas a at u + c: the fusion weights' ordering at a
 * temperature T (T = div of pct_weight_divisor gives the normalised fusion
 * weights themselves, nearly uniform for large sigma).  The student
 * distribution is q_c = softmax_c(-d_c / tau), d_c = |e_T(u) - e_a(u + c)|^2
 * the squared distance of the FP32 embeddings.  Loss = mean over anchors of
 * KL(p || q).  mu and inv_sd are not trained. */
typedef struct pct_train_opts {
    int steps;       /* Adam steps */
    int anchors;     /* anchors per step, drawn uniformly (interior, atlas) */
    float lr;        /* Adam learning rate (beta1 0.9, beta2 0.999, eps 1e-8) */
    float tau;       /* student temperature */
    float teacher_t; /* teacher temperature T; <= 0: div */
    uint64_t seed;   /* anchor sampling */
} pct_train_opts;

/* Loss and gradient (host, PCT_MLP_NPARAMS floats; 0 for mu, inv_sd and the
 * unused W1 rows) for the given anchors (host, [n_anchors][4] = z, x, y,
 * atlas).  Deterministic.  Synchronises. */
int pct_mlp_loss_grad_dev(const pct_problem* pb, const float* params, const int* anchors,
                          int n_anchors, float tau, float teacher_t, const float* d_img,
                          const float* d_atlas, double* loss, float* grad, pct_stream_t stream);

/* Trains `params` (host, updated in place) on the GPU; loss_trace (host,
 * `steps` values, may be NULL) receives each step's loss.  Synchronises. */
int pct_mlp_train_dev(const pct_problem* pb, float* params, const pct_train_opts* opts,
                      const float* d_img, const float* d_atlas, float* loss_trace,
                      pct_stream_t stream);

/* ---- kernel timing (no counterpart in the reference, which times the whole
 * wrapper with cudaEvents, cu:2152-2188) ---- */

/* While enabled, every fusion kernel launch (tiled or direct, and the
 * pruned kernel of pct_fuse_pruned_dev; not the cross-term pair pre-pass,
 * the finalize pass or the pruned path's item-filtered full search) is
 * bracketed by HIP events on its own stream.  Enabling clears the record.
 * At most 4096 launches are recorded. */
int pct_kernel_timing(int enable);

/* Waits for the recorded events; returns the summed kernel time and the
 * number of launches recorded since timing was enabled. */
int pct_kernel_timing_read(double* total_ms, int* launches);

/* ---- post-processing (cpp:1020-1106, 1221-1227, 1725-1739), host ---- */

/* NaN fill with the median of the 3x3x3 neighbourhood, iterated in place in
 * the reference's z, x, y order over the interior (cpp:1059-1106).  Returns
 * the number of passes, or a negative status.  Where a NaN voxel has no
 * non-NaN neighbour the reference reads uninitialised memory (cpp:1036-1041);
 * here it stays NaN for that pass, and the loop stops if a pass fills
 * nothing. */
int pct_regularize(int nh, int nx, int ny, int nz, float* arr);

/* round() every element (cpp:1221-1227). */
void pct_round(float* arr, size_t n);

/* v = round(4095.0f*(v/255.0f)) - 1024.0f (cpp:1733-1738). */
void pct_to_hu(float* arr, size_t n);

/* The whole chain of cpp:1725-1739: regularize, round, to HU. */
int pct_postprocess(int nh, int nx, int ny, int nz, float* arr);

/* Device versions of pct_regularize and pct_postprocess on a device volume,
 * enqueued on `stream`; bit-identical results.  The NaN fill visits the NaN
 * voxels in the reference's scan order, one at a time (cost grows with the
 * NaN count, not the volume).  Uses the library workspace.  Volumes of
 * nx*ny*nz >= 2^31 - 1 voxels return PCT_ERR_TOO_LARGE. */
int pct_regularize_dev(int nh, int nx, int ny, int nz, float* d_arr, pct_stream_t stream);
int pct_postprocess_dev(int nh, int nx, int ny, int nz, float* d_arr, pct_stream_t stream);

/* ---- on-disk formats and inputs ---- */

/* MatrixIO text format (cpp:252-320): lines num, rows, cols, depth, then the
 * values in memory order separated by blanks.  Allocates *data (free with
 * pct_free); dims4 = {rows, cols, depth, num} as the reference's
 * `dimensions` vector (cpp:268-271).  Values parse as the reference's
 * `iss >> float` does (correctly rounded), in parallel.  When a current
 * binary sidecar "<path>.pctbin" exists (see below), its values are read
 * instead of the text. */
int pct_matrixio_read(const char* path, float** data, int dims4[4]);

/* Writes "<path>.pctbin" beside the MatrixIO text file `path`: a header that
 * records the text file's size and modification time, then the values as raw
 * float32.  pct_matrixio_read uses it only while the text file keeps that
 * size and time.  No counterpart in the reference, whose inputs are
 * re-parsed on every run. */
int pct_matrixio_write_sidecar(const char* path);

/* Writes the MatrixIO text format with the reference's formatting (default
 * 6 significant digits, one blank after every value; cpp:289-320). */
int pct_matrixio_write(const char* path, const float* data, int num, int rows, int cols,
                       int depth);

/* Inputs as the reference's readMatFile (cpp:908-1018) fills its globals. */
typedef struct pct_volume_set {
    int nx, ny, nz, n_atlas;
    float* img;     /* nx*ny*nz */
    float* atlas;   /* n_atlas*nx*ny*nz */
    float* label;   /* n_atlas*nx*ny*nz, CT in HU (not yet normalised) */
} pct_volume_set;

/* Configuration file (cpp:908-960): line 1 a random flag; if set, lines
 * num, nx, ny, nz and uniform 10*U[0,1) data (cpp:961-973), drawn from a
 * seeded generator (the reference seeds rand() from the clock, cpp:918);
 * otherwise the atlas, label and image MatrixIO paths.  Buffers are
 * allocated; release with pct_volume_set_free. */
int pct_read_config(const char* path, uint64_t seed, pct_volume_set* out);
void pct_volume_set_free(pct_volume_set* vs);
void pct_free(void* p);

/* MAT-file (level 5) inputs of the reference's MATLAB build (ReadMatFile,
 * FindDimension, cpp:584-801) without libmat: struct `atlas` with fields
 * MRI1..MRIn, struct `atlas_label` with as many fields, and the target as the
 * remaining numeric variable, all nx x ny x nz; MATLAB's column-major order
 * is converted to the library layout (cpp:377-474).  Numeric classes are
 * converted to float; zlib-compressed elements are read.  Little-endian
 * files only.  PCT_ERR_FORMAT for anything else, PCT_ERR_DIMS when the
 * shapes or atlas counts differ.  Release with pct_volume_set_free. */
int pct_mat_read(const char* path, pct_volume_set* out);

/* Writes `data` (library layout) as a single-precision nx x ny x nz MATLAB
 * array named `name` to a MAT-file, as WriteOutputMatFile (cpp:835-906). */
int pct_mat_write(const char* path, const char* name, const float* data, int nx, int ny,
                  int nz);

/* Seeded uniform data of the reference's random mode (cpp:961-973):
 * every value 10*U[0,1). */
int pct_synth_uniform(int nx, int ny, int nz, int n_atlas, uint64_t seed,
                      float* img, float* atlas, float* label);

/* Seeded tissue-class phantom: N(0,1) noise, separable Gaussian blur
 * (sigma 3 voxels, taps to 4 sigma, mirror boundary), thresholds at the
 * 25/50/75 % quantiles -> classes {air, fat, soft tissue, bone};
 * MRI = {0, 0.8, 0.5, 0.2}[class] + N(0, 0.03),
 * CT  = {-1000, -100, 40, 1000}[class] + N(0, 20) HU.
 * Atlas a: if independent_anatomy, its own class map; otherwise the
 * target's class map shifted by a random integer vector in
 * [-max_shift, max_shift]^3 (edge-clamped).  Every atlas has its own noise.
 * atlas_ct is in HU (normalise with pct_normalize_labels). */
int pct_synth_tissue(int nx, int ny, int nz, int n_atlas, int max_shift,
                     int independent_anatomy, uint64_t seed,
                     float* target_mri, float* atlas_mri, float* atlas_ct);

/* ---- devices ---- */

/* Number of HIP devices (0 without a GPU; never fails). */
int pct_device_count(void);

/* Makes `device` current for the calling thread (cudaSetDevice(argv[7]) of
 * the reference, cu:1988); PCT_ERR_NO_DEVICE if it is not a gfx950 GPU. */
int pct_set_device(int device);

/* Releases the library's per-device workspaces and cached state. */
void pct_release(void);

#ifdef __cplusplus
}
#endif
#endif /* PCT_PCT_H */
