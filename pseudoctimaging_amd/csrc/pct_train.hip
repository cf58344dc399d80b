// Training of the voxel embedding MLP on gfx950, FP32 throughout.
//
// One step over A anchors with K = NH^3 window offsets each:
//   train_forward       one thread per voxel slot (A * (K+1) slots: the
//                       anchor's target voxel, then the atlas voxel at every
//                       offset): features, both hidden layers, embedding.
//                       Activations go to a row-major [row][slot] buffer so
//                       every store and later load is coalesced.
//   train_loss          one workgroup per anchor: teacher SSD and student
//                       distance for every offset, both log-softmaxes, the
//                       KL term, and the gradient with respect to every
//                       embedding of the anchor.
//   train_backward      one thread per slot: gradients of both hidden layers.
//   train_grad_partial  one workgroup per chunk of slots: every parameter's
//                       gradient summed over the chunk (LDS-staged tiles).
//   grad_reduce         partials summed in chunk order: the gradient does not
//                       depend on scheduling.
//   adam_step           Adam on the FP32 master parameters.
// No counterpart in the reference program.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pct_internal.h"
#include "pct_mlp.h"
#include "pct_tiled.h"

namespace pct {

// Activation buffer rows (each row holds one value for every slot).
constexpr int kRowX = 0;       // 30 features
constexpr int kRowOne = 30;    // constant 1 (bias gradients)
constexpr int kRowH1 = 31;     // 64 first hidden layer (after ReLU)
constexpr int kRowH2 = 95;     // 64 second hidden layer (after ReLU)
constexpr int kRowDE = 159;    // 16 d loss / d embedding
constexpr int kRowDH2 = 175;   // 64 d loss / d second layer pre-activation
constexpr int kRowDH1 = 239;   // 64 d loss / d first layer pre-activation
constexpr int kRowsGrad = 303; // rows the gradient pass reads
constexpr int kRowE = 303;     // 16 embedding
constexpr int kRows = 319;

constexpr int kTrainThreads = 256;
constexpr int kGradTile = 32;  // slots per LDS tile of the gradient pass
constexpr int kParamsPerThread = (kNetParams + kTrainThreads - 1) / kTrainThreads;
constexpr int kMaxWindow = 15; // NH

struct TrainArgs {
    const float* img;
    const float* atlas;       // [n][N]
    long long N;
    int nx, ny, nz, hw;
    const int* anchors;       // [A][4] z, x, y, atlas
    int n_anchors;
    int K;                    // NH^3
    long long V;              // A * (K + 1)
    const float* params;      // device copy of the parameter vector
    float* buf;               // [kRows][V]
    float* loss;              // [A]
    float mu, inv_sd;
    float teacher_t, tau;
};

__device__ __forceinline__ void slot_voxel(const TrainArgs& a, long long v, int* z, int* x, int* y,
                                           const float** vol) {
    const int i = (int)(v / (a.K + 1));
    const int k = (int)(v - (long long)i * (a.K + 1));
    const int* an = a.anchors + 4 * i;
    *z = an[0];
    *x = an[1];
    *y = an[2];
    *vol = a.img;
    if (k > 0) {
        const int c = k - 1, nw = 2 * a.hw + 1;
        *z += c / (nw * nw) - a.hw;
        *x += (c / nw) % nw - a.hw;
        *y += c % nw - a.hw;
        *vol = a.atlas + (long long)an[3] * a.N;
    }
}

__global__ __launch_bounds__(kTrainThreads) void train_forward(TrainArgs a) {
    __shared__ float w[kNetParams];
    for (int i = threadIdx.x; i < kNetParams; i += kTrainThreads) w[i] = a.params[i];
    __syncthreads();
    const long long v = (long long)blockIdx.x * kTrainThreads + threadIdx.x;
    if (v >= a.V) return;
    int z, x, y;
    const float* vol;
    slot_voxel(a, v, &z, &x, &y, &vol);
    float* __restrict__ B = a.buf;
    const long long V = a.V;
    float f[kFeat];
#pragma unroll
    for (int k = 0; k < kFeat; ++k) {
        f[k] = feature(vol, a.nx, a.ny, a.nz, z, x, y, k, a.mu, a.inv_sd);
        B[(kRowX + k) * V + v] = f[k];
    }
    B[kRowOne * V + v] = 1.f;
    float h1[kHid];
#pragma unroll
    for (int m = 0; m < kHid; ++m) {
        float acc = w[kOffB1 + m];
#pragma unroll
        for (int k = 0; k < kFeat; ++k) acc = fmaf(f[k], w[kOffW1 + k * kHid + m], acc);
        h1[m] = fmaxf(acc, 0.f);
        B[(kRowH1 + m) * V + v] = h1[m];
    }
    float e[kEmb];
#pragma unroll
    for (int j = 0; j < kEmb; ++j) e[j] = w[kOffB3 + j];
#pragma unroll
    for (int mc = 0; mc < kHid; mc += 16) {
        float acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = w[kOffB2 + mc + j];
#pragma unroll
        for (int k = 0; k < kHid; ++k)
#pragma unroll
            for (int j = 0; j < 16; ++j)
                acc[j] = fmaf(h1[k], w[kOffW2 + k * kHid + mc + j], acc[j]);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float h = fmaxf(acc[j], 0.f);
            B[(kRowH2 + mc + j) * V + v] = h;
#pragma unroll
            for (int q = 0; q < kEmb; ++q) e[q] = fmaf(h, w[kOffW3 + (mc + j) * kEmb + q], e[q]);
        }
    }
#pragma unroll
    for (int j = 0; j < kEmb; ++j) B[(kRowE + j) * V + v] = e[j];
}

// Deterministic workgroup reductions (fixed pairing over kTrainThreads).
__device__ __forceinline__ float block_sum(float v, float* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kTrainThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const float r = red[0];
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_max(float v, float* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kTrainThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const float r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kTrainThreads) void train_loss(TrainArgs a) {
    extern __shared__ float sh[];
    float* lp = sh;                       // [K] teacher logits
    float* lq = sh + a.K;                 // [K] student logits
    float* red = sh + 2 * a.K;            // [kTrainThreads]
    float* eT = red + kTrainThreads;      // [16]
    float* tp = eT + kEmb;                // [27] target patch
    const int i = blockIdx.x;
    const long long base = (long long)i * (a.K + 1);
    const long long V = a.V;
    float* __restrict__ B = a.buf;
    const int* an = a.anchors + 4 * i;
    const int z = an[0], x = an[1], y = an[2];
    const float* __restrict__ A = a.atlas + (long long)an[3] * a.N;
    const long long plane = (long long)a.nx * a.ny;
    if (threadIdx.x < kEmb) eT[threadIdx.x] = B[(kRowE + threadIdx.x) * V + base];
    if (threadIdx.x < 27) {
        const int p = threadIdx.x;
        tp[p] = a.img[(z + p / 9 - 1) * plane + (long long)(x + (p / 3) % 3 - 1) * a.ny +
                      (y + p % 3 - 1)];
    }
    __syncthreads();
    const int nw = 2 * a.hw + 1;
    float mp = -__builtin_huge_valf(), mq = -__builtin_huge_valf();
    for (int c = threadIdx.x; c < a.K; c += kTrainThreads) {
        const int zc = z + c / (nw * nw) - a.hw, xc = x + (c / nw) % nw - a.hw,
                  yc = y + c % nw - a.hw;
        float ssd = 0.f;
        for (int p = 0; p < 27; ++p) {
            const float d = tp[p] - A[(zc + p / 9 - 1) * plane +
                                      (long long)(xc + (p / 3) % 3 - 1) * a.ny + (yc + p % 3 - 1)];
            ssd = fmaf(d, d, ssd);
        }
        float dist = 0.f;
#pragma unroll
        for (int j = 0; j < kEmb; ++j) {
            const float d = eT[j] - B[(kRowE + j) * V + base + 1 + c];
            dist = fmaf(d, d, dist);
        }
        lp[c] = -ssd / a.teacher_t;
        lq[c] = -dist / a.tau;
        mp = fmaxf(mp, lp[c]);
        mq = fmaxf(mq, lq[c]);
    }
    mp = block_max(mp, red);
    mq = block_max(mq, red);
    float zp = 0.f, zq = 0.f;
    for (int c = threadIdx.x; c < a.K; c += kTrainThreads) {
        zp += expf(lp[c] - mp);
        zq += expf(lq[c] - mq);
    }
    zp = block_sum(zp, red);
    zq = block_sum(zq, red);
    const float lzp = mp + logf(zp), lzq = mq + logf(zq);
    const float gscale = 1.f / (a.tau * (float)a.n_anchors);
    float kl = 0.f;
    float gT[kEmb];
#pragma unroll
    for (int j = 0; j < kEmb; ++j) gT[j] = 0.f;
    for (int c = threadIdx.x; c < a.K; c += kTrainThreads) {
        const float lpn = lp[c] - lzp, lqn = lq[c] - lzq;
        const float p = expf(lpn), q = expf(lqn);
        kl = fmaf(p, lpn - lqn, kl);
        const float g = (p - q) * gscale;   // d loss / d dist_c
#pragma unroll
        for (int j = 0; j < kEmb; ++j) {
            const long long idx = (kRowE + j) * V + base + 1 + c;
            const float diff = eT[j] - B[idx];
            gT[j] = fmaf(2.f * g, diff, gT[j]);
            B[(kRowDE + j) * V + base + 1 + c] = -2.f * g * diff;
        }
    }
    kl = block_sum(kl, red);
#pragma unroll
    for (int j = 0; j < kEmb; ++j) {
        const float s = block_sum(gT[j], red);
        if (threadIdx.x == 0) B[(kRowDE + j) * V + base] = s;
    }
    if (threadIdx.x == 0) a.loss[i] = kl / (float)a.n_anchors;
}

__global__ __launch_bounds__(kTrainThreads) void train_backward(TrainArgs a) {
    __shared__ float w2[kHid * kHid];
    __shared__ float w3[kHid * kEmb];
    for (int i = threadIdx.x; i < kHid * kHid; i += kTrainThreads) w2[i] = a.params[kOffW2 + i];
    for (int i = threadIdx.x; i < kHid * kEmb; i += kTrainThreads) w3[i] = a.params[kOffW3 + i];
    __syncthreads();
    const long long v = (long long)blockIdx.x * kTrainThreads + threadIdx.x;
    if (v >= a.V) return;
    float* __restrict__ B = a.buf;
    const long long V = a.V;
    float de[kEmb];
#pragma unroll
    for (int j = 0; j < kEmb; ++j) de[j] = B[(kRowDE + j) * V + v];
    float dh2[kHid];
#pragma unroll
    for (int m = 0; m < kHid; ++m) {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < kEmb; ++j) s = fmaf(w3[m * kEmb + j], de[j], s);
        dh2[m] = B[(kRowH2 + m) * V + v] > 0.f ? s : 0.f;
        B[(kRowDH2 + m) * V + v] = dh2[m];
    }
#pragma unroll 4
    for (int k = 0; k < kHid; ++k) {
        float s = 0.f;
#pragma unroll
        for (int m = 0; m < kHid; ++m) s = fmaf(w2[k * kHid + m], dh2[m], s);
        B[(kRowDH1 + k) * V + v] = B[(kRowH1 + k) * V + v] > 0.f ? s : 0.f;
    }
}

// Rows whose product, summed over slots, is the gradient of parameter p
// (ra = -1: no gradient).
__device__ __forceinline__ void param_rows(int p, int* ra, int* rb) {
    *ra = -1;
    *rb = 0;
    if (p < kOffB1) {
        const int k = p / kHid, m = p % kHid;
        if (k < kFeat) {
            *ra = kRowX + k;
            *rb = kRowDH1 + m;
        }
    } else if (p < kOffW2) {
        *ra = kRowOne;
        *rb = kRowDH1 + (p - kOffB1);
    } else if (p < kOffB2) {
        const int q = p - kOffW2;
        *ra = kRowH1 + q / kHid;
        *rb = kRowDH2 + q % kHid;
    } else if (p < kOffW3) {
        *ra = kRowOne;
        *rb = kRowDH2 + (p - kOffB2);
    } else if (p < kOffB3) {
        const int q = p - kOffW3;
        *ra = kRowH2 + q / kEmb;
        *rb = kRowDE + q % kEmb;
    } else if (p < kNetParams) {
        *ra = kRowOne;
        *rb = kRowDE + (p - kOffB3);
    }
}

__global__ __launch_bounds__(kTrainThreads) void train_grad_partial(TrainArgs a, float* partial,
                                                                    long long chunk) {
    __shared__ float t[kRowsGrad][kGradTile + 1];
    const long long v0 = (long long)blockIdx.x * chunk;
    const long long v1 = min(a.V, v0 + chunk);
    int ra[kParamsPerThread], rb[kParamsPerThread];
    float acc[kParamsPerThread];
#pragma unroll
    for (int q = 0; q < kParamsPerThread; ++q) {
        const int p = threadIdx.x + q * kTrainThreads;
        param_rows(p < kNetParams ? p : kNetParams, &ra[q], &rb[q]);
        acc[q] = 0.f;
    }
    const float* __restrict__ B = a.buf;
    for (long long vt = v0; vt < v1; vt += kGradTile) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < kRowsGrad * kGradTile; idx += kTrainThreads) {
            const int row = idx / kGradTile, col = idx % kGradTile;
            const long long vv = vt + col;
            t[row][col] = vv < v1 ? B[row * a.V + vv] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kParamsPerThread; ++q) {
            if (ra[q] >= 0) {
                float s = acc[q];
                for (int col = 0; col < kGradTile; ++col) s = fmaf(t[ra[q]][col], t[rb[q]][col], s);
                acc[q] = s;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kParamsPerThread; ++q) {
        const int p = threadIdx.x + q * kTrainThreads;
        if (p < kNetParams) partial[(long long)blockIdx.x * kNetParams + p] = acc[q];
    }
}

__global__ void grad_reduce(const float* __restrict__ partial, int nblk, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PCT_MLP_NPARAMS) return;
    float s = 0.f;
    if (p < kNetParams)
        for (int b = 0; b < nblk; ++b) s += partial[(long long)b * kNetParams + p];
    grad[p] = s;
}

__global__ void adam_step(float* __restrict__ params, float* __restrict__ m1,
                          float* __restrict__ m2, const float* __restrict__ grad, float lr,
                          float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNetParams) return;
    const float g = grad[p];
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float b = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
}

// Device buffers of one training run, carved from the library workspace.
struct TrainPlan {
    int K = 0, n_anchors = 0, nblk = 0;
    long long V = 0, chunk = 0;
    size_t o_params = 0, o_m1 = 0, o_m2 = 0, o_grad = 0, o_anchors = 0, o_buf = 0, o_loss = 0,
           o_partial = 0, bytes = 0;
};

static TrainPlan plan_training(int nh, int n_anchors) {
    TrainPlan t;
    t.K = nh * nh * nh;
    t.n_anchors = n_anchors;
    t.V = (long long)n_anchors * (t.K + 1);
    t.nblk = (int)std::min<long long>(1024, (t.V + 511) / 512);
    t.chunk = (t.V + t.nblk - 1) / t.nblk;
    t.chunk = (t.chunk + kGradTile - 1) / kGradTile * kGradTile;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off = 0;
    t.o_params = off; off = align(off + 4 * PCT_MLP_NPARAMS);
    t.o_m1 = off; off = align(off + 4 * PCT_MLP_NPARAMS);
    t.o_m2 = off; off = align(off + 4 * PCT_MLP_NPARAMS);
    t.o_grad = off; off = align(off + 4 * PCT_MLP_NPARAMS);
    t.o_anchors = off; off = align(off + 16 * (size_t)n_anchors);
    t.o_buf = off; off = align(off + 4 * (size_t)kRows * t.V);
    t.o_loss = off; off = align(off + 4 * (size_t)n_anchors);
    t.o_partial = off; off = align(off + 4 * (size_t)t.nblk * kNetParams);
    t.bytes = off;
    return t;
}

static int check_train_problem(const pct_problem* pb) {
    int rc = validate_problem(pb);
    if (rc) return rc;
    if (pb->nh > kMaxWindow || pb->patch != 3) return PCT_ERR_UNSUPPORTED;
    const int r = (pb->nh - 1) / 2 + 1;
    if (pb->nx <= 2 * r || pb->ny <= 2 * r || pb->nz <= 2 * r) return PCT_ERR_DIMS;
    return PCT_OK;
}

// Enqueues forward, loss, backward and the gradient into `grad` (device).
static int enqueue_step(const TrainPlan& t, const TrainArgs& a, char* w, hipStream_t s) {
    const unsigned vb = (unsigned)((t.V + kTrainThreads - 1) / kTrainThreads);
    hipLaunchKernelGGL(train_forward, dim3(vb), dim3(kTrainThreads), 0, s, a);
    const size_t loss_lds = sizeof(float) * (2 * (size_t)t.K + kTrainThreads + kEmb + 27);
    hipLaunchKernelGGL(train_loss, dim3(t.n_anchors), dim3(kTrainThreads), loss_lds, s, a);
    hipLaunchKernelGGL(train_backward, dim3(vb), dim3(kTrainThreads), 0, s, a);
    float* partial = reinterpret_cast<float*>(w + t.o_partial);
    hipLaunchKernelGGL(train_grad_partial, dim3(t.nblk), dim3(kTrainThreads), 0, s, a, partial,
                       t.chunk);
    hipLaunchKernelGGL(grad_reduce, dim3((PCT_MLP_NPARAMS + 255) / 256), dim3(256), 0, s,
                       partial, t.nblk, reinterpret_cast<float*>(w + t.o_grad));
    return hipGetLastError() == hipSuccess ? PCT_OK : PCT_ERR_HIP;
}

static TrainArgs make_train_args(const pct_problem* pb, const TrainPlan& t, const float* params,
                                 float tau, float teacher_t, const float* d_img,
                                 const float* d_atlas, char* w) {
    TrainArgs a;
    a.img = d_img;
    a.atlas = d_atlas;
    a.N = (long long)pb->nx * pb->ny * pb->nz;
    a.nx = pb->nx;
    a.ny = pb->ny;
    a.nz = pb->nz;
    a.hw = (pb->nh - 1) / 2;
    a.anchors = reinterpret_cast<const int*>(w + t.o_anchors);
    a.n_anchors = t.n_anchors;
    a.K = t.K;
    a.V = t.V;
    a.params = reinterpret_cast<const float*>(w + t.o_params);
    a.buf = reinterpret_cast<float*>(w + t.o_buf);
    a.loss = reinterpret_cast<float*>(w + t.o_loss);
    a.mu = params[kOffMu];
    a.inv_sd = params[kOffSd];
    a.teacher_t = teacher_t > 0.f ? teacher_t : (float)problem_divisor(pb);
    a.tau = tau;
    return a;
}

}  // namespace pct

using namespace pct;

extern "C" int pct_mlp_loss_grad_dev(const pct_problem* pb, const float* params,
                                     const int* anchors, int n_anchors, float tau,
                                     float teacher_t, const float* d_img, const float* d_atlas,
                                     double* loss, float* grad, pct_stream_t stream) {
    int rc = check_train_problem(pb);
    if (rc) return rc;
    if (!params || !anchors || n_anchors < 1 || !(tau > 0.f) || !d_img || !d_atlas || !loss ||
        !grad)
        return PCT_ERR_INVALID;
    const int r = (pb->nh - 1) / 2 + 1;
    for (int i = 0; i < n_anchors; ++i) {
        const int* an = anchors + 4 * i;
        if (an[0] < r || an[0] >= pb->nz - r || an[1] < r || an[1] >= pb->nx - r || an[2] < r ||
            an[2] >= pb->ny - r || an[3] < 0 || an[3] >= pb->n_atlas)
            return PCT_ERR_INVALID;
    }
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const TrainPlan t = plan_training(pb->nh, n_anchors);
    void* ws = nullptr;
    rc = ensure_workspace(dev, t.bytes, &ws);
    if (rc) return rc;
    char* w = static_cast<char*>(ws);
    if (hipMemcpyAsync(w + t.o_params, params, 4 * PCT_MLP_NPARAMS, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(w + t.o_anchors, anchors, 16 * (size_t)n_anchors, hipMemcpyHostToDevice,
                       s) != hipSuccess)
        return PCT_ERR_HIP;
    const TrainArgs a = make_train_args(pb, t, params, tau, teacher_t, d_img, d_atlas, w);
    rc = enqueue_step(t, a, w, s);
    if (rc) return rc;
    std::vector<float> lv(n_anchors);
    if (hipMemcpyAsync(grad, w + t.o_grad, 4 * PCT_MLP_NPARAMS, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(lv.data(), w + t.o_loss, 4 * (size_t)n_anchors, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return PCT_ERR_HIP;
    double sum = 0.0;
    for (float v : lv) sum += v;
    *loss = sum;
    return PCT_OK;
}

extern "C" int pct_mlp_train_dev(const pct_problem* pb, float* params,
                                 const pct_train_opts* opts, const float* d_img,
                                 const float* d_atlas, float* loss_trace, pct_stream_t stream) {
    int rc = check_train_problem(pb);
    if (rc) return rc;
    if (!params || !opts || opts->steps < 0 || opts->anchors < 1 || !(opts->tau > 0.f) ||
        !(opts->lr > 0.f) || !d_img || !d_atlas)
        return PCT_ERR_INVALID;
    int dev = 0;
    rc = check_device(&dev);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const TrainPlan t = plan_training(pb->nh, opts->anchors);
    void* ws = nullptr;
    rc = ensure_workspace(dev, t.bytes, &ws);
    if (rc) return rc;
    char* w = static_cast<char*>(ws);
    float* d_params = reinterpret_cast<float*>(w + t.o_params);
    if (hipMemcpyAsync(d_params, params, 4 * PCT_MLP_NPARAMS, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemsetAsync(w + t.o_m1, 0, 4 * PCT_MLP_NPARAMS, s) != hipSuccess ||
        hipMemsetAsync(w + t.o_m2, 0, 4 * PCT_MLP_NPARAMS, s) != hipSuccess)
        return PCT_ERR_HIP;
    const TrainArgs a =
        make_train_args(pb, t, params, opts->tau, opts->teacher_t, d_img, d_atlas, w);
    const int r = (pb->nh - 1) / 2 + 1;
    const int iz = pb->nz - 2 * r, ix = pb->nx - 2 * r, iy = pb->ny - 2 * r;
    uint64_t st = opts->seed ^ 0x9E3779B97F4A7C15ULL;
    std::vector<int> anc(4 * (size_t)opts->anchors);
    std::vector<float> lv(opts->anchors);
    double b1 = 1.0, b2 = 1.0;
    for (int step = 0; step < opts->steps; ++step) {
        for (int i = 0; i < opts->anchors; ++i) {
            anc[4 * i + 0] = r + (int)(splitmix64(st) % (uint64_t)iz);
            anc[4 * i + 1] = r + (int)(splitmix64(st) % (uint64_t)ix);
            anc[4 * i + 2] = r + (int)(splitmix64(st) % (uint64_t)iy);
            anc[4 * i + 3] = (int)(splitmix64(st) % (uint64_t)pb->n_atlas);
        }
        if (hipMemcpyAsync(w + t.o_anchors, anc.data(), 16 * (size_t)opts->anchors,
                           hipMemcpyHostToDevice, s) != hipSuccess)
            return PCT_ERR_HIP;
        rc = enqueue_step(t, a, w, s);
        if (rc) return rc;
        b1 *= 0.9;
        b2 *= 0.999;
        hipLaunchKernelGGL(adam_step, dim3((kNetParams + 255) / 256), dim3(256), 0, s, d_params,
                           reinterpret_cast<float*>(w + t.o_m1),
                           reinterpret_cast<float*>(w + t.o_m2),
                           reinterpret_cast<const float*>(w + t.o_grad), opts->lr,
                           (float)(1.0 - b1), (float)(1.0 - b2));
        if (hipGetLastError() != hipSuccess) return PCT_ERR_HIP;
        // the anchor vector is rewritten next step and the loss is read now
        if (hipMemcpyAsync(lv.data(), w + t.o_loss, 4 * (size_t)opts->anchors,
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return PCT_ERR_HIP;
        if (loss_trace) {
            double sum = 0.0;
            for (float v : lv) sum += v;
            loss_trace[step] = (float)sum;
        }
    }
    if (hipMemcpyAsync(params, d_params, 4 * kNetParams, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return PCT_ERR_HIP;
    return PCT_OK;
}
