/* CPU oracle: plain-C restatement of the reference CPU program's fusion path
 * (src/CPU/NewPatchedBrainPureC.cpp).  TEST INFRASTRUCTURE ONLY - see
 * pct_oracle.h for who may use it and what pins each function. */
#include "pct_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

void oracle_normalize_labels(float* label, size_t n) {
    /* cpp:1586-1588, float32 in this operation order */
    for (size_t i = 0; i < n; ++i) label[i] = 255.0f * ((label[i] + 1024.0f) / 4095.0f);
}

double oracle_div(double sigma, double beta) {
    /* cpp:1601: (2 * 27 * beta*pow(sigma, 2)) evaluated left to right */
    return (2 * 27 * beta * pow(sigma, 2));
}

/* PatchSimilarity (cpp:1184-1212): 27-term SSD in float64 over the target
 * patch at (z,x,y) and the atlas patch at (w,u,v), loop order z, x, y
 * (cpp:1198-1206), then exp(-SSD/div) (cpp:1208-1210). */
static inline double patch_weight(int nx, int ny, const float* img, const float* atl,
                                  int z, int x, int y, int w, int u, int v, double div) {
    const long pl = (long)nx * ny;
    double sum = 0;
    for (int k = -1; k <= 1; ++k)
        for (int i = -1; i <= 1; ++i)
            for (int j = -1; j <= 1; ++j) {
                double p = img[(z + k) * pl + (long)(x + i) * ny + (y + j)];
                double pa = atl[(w + k) * pl + (long)(u + i) * ny + (v + j)];
                sum += pow(p - pa, 2);
            }
    double arg = sum / div;
    return exp(-arg);
}

/* Accumulation for one voxel, loop order atlas, w (z), u (x), v (y)
 * (cpp:1399-1432). */
static inline void voxel_sums(int nx, int ny, int nz, int num, int hw,
                              const float* img, const float* atlas, const float* lab,
                              double div, int z, int x, int y, double* P, double* S) {
    const long pitch = (long)nx * ny * nz;
    const long pl = (long)nx * ny;
    double product = 0, sum_weight = 0;
    for (int at = 0; at < num; ++at) {
        const float* atl = atlas + pitch * at;
        const float* lb = lab + pitch * at;
        for (int w = z - hw; w <= z + hw; ++w)
            for (int u = x - hw; u <= x + hw; ++u)
                for (int v = y - hw; v <= y + hw; ++v) {
                    double label = lb[w * pl + (long)u * ny + v];
                    double weight = patch_weight(nx, ny, img, atl, z, x, y, w, u, v, div);
                    product += weight * label;
                    sum_weight += weight;
                }
    }
    *P = product;
    *S = sum_weight;
}

void oracle_fuse(int nx, int ny, int nz, int num, int nh,
                 const float* img, const float* atlas, const float* label_norm,
                 float* out, double div, int threads) {
    const int hp = 1;                 /* patch 3x3x3, cpp:1240-1242 */
    const int hw = (nh - 1) / 2;      /* cpp:1243 */
    const long pl = (long)nx * ny;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int z = hw + hp; z < nz - hw - hp; ++z)          /* cpp:1374-1377 */
        for (int x = hw + hp; x < nx - hw - hp; ++x)
            for (int y = hw + hp; y < ny - hw - hp; ++y) {
                double P, S;
                voxel_sums(nx, ny, nz, num, hw, img, atlas, label_norm, div, z, x, y, &P, &S);
                /* cpp:1437-1443 */
                if (fabs(P) > DBL_EPSILON && fabs(S) > DBL_EPSILON)
                    out[z * pl + (long)x * ny + y] = (float)(P / S);
                else
                    out[z * pl + (long)x * ny + y] = NAN;
            }
}

double oracle_voxel(int nx, int ny, int nz, int num, int nh,
                    const float* img, const float* atlas, const float* label_norm,
                    double div, int z, int x, int y, double* P, double* S) {
    const int hw = (nh - 1) / 2;
    double p, s;
    voxel_sums(nx, ny, nz, num, hw, img, atlas, label_norm, div, z, x, y, &p, &s);
    if (P) *P = p;
    if (S) *S = s;
    if (fabs(p) > DBL_EPSILON && fabs(s) > DBL_EPSILON) return p / s;
    return NAN;
}

void oracle_voxels(int nx, int ny, int nz, int num, int nh,
                   const float* img, const float* atlas, const float* label_norm,
                   double div, const long* idx, long count, double* ratio_out,
                   int threads) {
    const long pl = (long)nx * ny;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
    for (long k = 0; k < count; ++k) {
        long id = idx[k];
        int z = (int)(id / pl);
        int x = (int)((id % pl) / ny);
        int y = (int)(id % ny);
        ratio_out[k] = oracle_voxel(nx, ny, nz, num, nh, img, atlas, label_norm, div,
                                    z, x, y, NULL, NULL);
    }
}

double oracle_median(const float* patch, int length) {
    /* cpp:1020-1043: keep the non-NaN entries, stable sort, middle value or
     * mean of the two middle values in float64. */
    float v[64];
    int cnt = 0;
    for (int j = 0; j < length && j < 64; ++j)
        if (!isnan(patch[j])) v[cnt++] = patch[j];
    if (cnt <= 0) return NAN;
    for (int i = 1; i < cnt; ++i) {          /* stable insertion sort */
        float key = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > key) { v[j + 1] = v[j]; --j; }
        v[j + 1] = key;
    }
    if (cnt % 2 == 0) {
        double a = v[cnt / 2 - 1];
        double b = v[cnt / 2];
        return (a + b) / 2;
    }
    return (double)v[cnt / 2];
}

int oracle_regularize(int hw, int hp, int nx, int ny, int nz, float* arr) {
    const long l1 = (long)nx * ny;
    float patch[27];
    int passes = 0;
    int nan_flag = 1;
    while (nan_flag) {                        /* cpp:1066-1102 */
        nan_flag = 0;
        long filled = 0;
        passes++;
        for (int z = hw + hp; z < nz - hw - hp; ++z)
            for (int x = hw + hp; x < nx - hw - hp; ++x)
                for (int y = hw + hp; y < ny - hw - hp; ++y) {
                    long id = z * l1 + (long)x * ny + y;
                    if (!isnan(arr[id])) continue;
                    int cnt = 0;
                    for (int r = z - hp; r <= z + hp; ++r)
                        for (int s = x - hp; s <= x + hp; ++s)
                            for (int t = y - hp; t <= y + hp; ++t)
                                patch[cnt++] = arr[r * l1 + (long)s * ny + t];
                    arr[id] = (float)oracle_median(patch, 27);
                    if (isnan(arr[id])) nan_flag = 1;
                    else filled++;
                }
        if (nan_flag && filled == 0) break;   /* divergence, see header */
    }
    return passes;
}

void oracle_round(float* arr, size_t n) {
    for (size_t i = 0; i < n; ++i) arr[i] = roundf(arr[i]);   /* cpp:1224 */
}

void oracle_to_hu(float* arr, size_t n) {
    const float max_labels = 4095.0f, min_labels = 1024.0f;    /* cpp:1733-1734 */
    for (size_t i = 0; i < n; ++i)
        arr[i] = roundf(max_labels * (arr[i] / 255.0f)) - min_labels;  /* cpp:1737 */
}

void oracle_postprocess(int nh, int nx, int ny, int nz, float* arr) {
    size_t n = (size_t)nx * ny * nz;
    oracle_regularize((nh - 1) / 2, 1, nx, ny, nz, arr);
    oracle_round(arr, n);
    oracle_to_hu(arr, n);
}
