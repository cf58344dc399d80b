// The voxel embedding MLP shared by inference (pct_learned.hip) and training
// (pct_train.hip): parameter layout, input features, BF16 rounding.
#pragma once

#include <hip/hip_runtime.h>

#include "pct/pct.h"

namespace pct {

// Parameter vector layout: W1[32][64] (rows 30, 31 unused), b1[64],
// W2[64][64], b2[64], W3[64][16], b3[16], mu, inv_sd.  W is input-major.
constexpr int kOffW1 = 0, kOffB1 = 2048, kOffW2 = 2112, kOffB2 = 6208, kOffW3 = 6272,
              kOffB3 = 7296, kOffMu = 7312, kOffSd = 7313;
constexpr int kNetParams = 7312;   // trained parameters (everything but mu, inv_sd)
constexpr int kFeat = 30, kHid = 64, kEmb = 16;
static_assert(kOffSd + 1 == PCT_MLP_NPARAMS, "parameter layout");

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
    return __uint_as_float((unsigned)h << 16);
}

// g(u): the normalised value inside the volume, 0 outside.
__device__ __forceinline__ float gval(const float* __restrict__ v, int nx, int ny, int nz, int z,
                                      int x, int y, float mu, float inv_sd) {
    if (z < 0 || z >= nz || x < 0 || x >= nx || y < 0 || y >= ny) return 0.f;
    return (v[((long long)z * nx + x) * ny + y] - mu) * inv_sd;
}

// Feature k of voxel (z, x, y): k < 27 the 3x3x3 neighbourhood
// f[9 pz + 3 px + py] = g(u + (pz-1, px-1, py-1)); k = 27, 28, 29 the
// central differences (g(u+e) - g(u-e)) / 2 along z, x, y; 0 beyond.
__device__ __forceinline__ float feature(const float* __restrict__ v, int nx, int ny, int nz,
                                         int z, int x, int y, int k, float mu, float inv_sd) {
    if (k < 27)
        return gval(v, nx, ny, nz, z + k / 9 - 1, x + (k / 3) % 3 - 1, y + k % 3 - 1, mu, inv_sd);
    if (k == 27)
        return 0.5f * (gval(v, nx, ny, nz, z + 1, x, y, mu, inv_sd) -
                       gval(v, nx, ny, nz, z - 1, x, y, mu, inv_sd));
    if (k == 28)
        return 0.5f * (gval(v, nx, ny, nz, z, x + 1, y, mu, inv_sd) -
                       gval(v, nx, ny, nz, z, x - 1, y, mu, inv_sd));
    if (k == 29)
        return 0.5f * (gval(v, nx, ny, nz, z, x, y + 1, mu, inv_sd) -
                       gval(v, nx, ny, nz, z, x, y - 1, mu, inv_sd));
    return 0.f;
}

}  // namespace pct
