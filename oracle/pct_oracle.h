/* CPU oracle for the multi-atlas patch-fusion path: a plain-C restatement of
 * the reference CPU program (src/CPU/NewPatchedBrainPureC.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (pseudoctimaging_amd/, libpct.so) never links or calls it.
 *
 * Volumes are float32 with linear index z*nx*ny + x*ny + y (y fastest);
 * atlas a starts at a*nx*ny*nz (cpp:1288).
 *
 * Parity pins (tests/test_oracle.py):
 *   oracle_fuse / oracle_voxel   - bitwise vs the compiled reference
 *                                  (oracle/_ref) on golden fixtures, plus
 *                                  closed forms and invariants.
 *   oracle_median / regularize / round / to_hu - bitwise vs the compiled
 *                                  reference on NaN-seeded fields.
 *   oracle_normalize_labels / oracle_div - closed-form values.
 * libm boundary (glibc exp/pow) is shared with the compiled reference and
 * is the one part no check pins across machines.
 */
#ifndef PCT_ORACLE_H
#define PCT_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* cpp:1583-1593: label'[i] = 255.0f*((label[i]+1024.0f)/4095.0f), float32. */
void oracle_normalize_labels(float* label, size_t n);

/* cpp:1601: div = 2*27*beta*sigma^2, float64. */
double oracle_div(double sigma, double beta);

/* cpp:1350-1449 (OpenMP over z, cpp:1374) / cpp:1228-1337 (threads <= 1).
 * Writes only interior voxels r <= z < nz-r (same for x, y), r = (nh-1)/2+1;
 * everything else keeps the caller's value. */
void oracle_fuse(int nx, int ny, int nz, int num, int nh,
                 const float* img, const float* atlas, const float* label_norm,
                 float* out, double div, int threads);

/* One interior voxel of oracle_fuse, returning the float64 sums P, S and the
 * float64 ratio P/S (NaN under the rule of cpp:1437-1443). */
double oracle_voxel(int nx, int ny, int nz, int num, int nh,
                    const float* img, const float* atlas, const float* label_norm,
                    double div, int z, int x, int y, double* P, double* S);

/* Same as oracle_voxel for a list of voxels (flat indices), parallel. */
void oracle_voxels(int nx, int ny, int nz, int num, int nh,
                   const float* img, const float* atlas, const float* label_norm,
                   double div, const long* idx, long count, double* ratio_out,
                   int threads);

/* cpp:1020-1043.  Divergence: with no non-NaN entry the reference reads
 * uninitialised memory; the oracle returns NaN. */
double oracle_median(const float* patch, int length);

/* cpp:1059-1106, in-place Gauss-Seidel NaN fill over the interior.
 * Divergence: stops when a whole pass fills nothing (the reference would
 * loop forever or read uninitialised memory).  Returns the number of
 * passes. */
int oracle_regularize(int half_window, int half_patch, int nx, int ny, int nz, float* arr);

/* cpp:1221-1227 */
void oracle_round(float* arr, size_t n);

/* cpp:1733-1738: v = round(4095.0f*(v/255.0f)) - 1024.0f */
void oracle_to_hu(float* arr, size_t n);

/* cpp:1725-1738: regularize, round, to_hu. */
void oracle_postprocess(int nh, int nx, int ny, int nz, float* arr);

#ifdef __cplusplus
}
#endif
#endif
