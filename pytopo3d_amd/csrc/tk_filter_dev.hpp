// Device helpers of the conical density filter and the Heaviside projection.
//
// The filter matrix H of pytopo3d/utils/filter.py:122-228 is a translation-
// invariant stencil truncated at the domain boundary: weight rmin - |d| for
// every integer offset d with rmin - |d| > 1e-9.  It is never stored; the
// stencil table (offsets + weights, at most kMaxStencil entries) is staged in
// LDS and applied as a gather.  H is symmetric, so H^T = H.
#pragma once

#include "tk_common.hpp"

namespace tk {

__device__ __forceinline__ void decode_offset(int o, int& dy, int& dx, int& dz) {
  dy = (o & 255) - 16;
  dx = ((o >> 8) & 255) - 16;
  dz = ((o >> 16) & 255) - 16;
}

// OC candidate of one element at multiplier lam (pytopo3d/utils/oc_update.py:
// 25-124): clip(max(x - m, min(x + m, x sqrt(max(ratio / (lam + 1e-30),
// 1e-30)))), 0, 1) on free elements, 1 / 0 on protected solid / void.  Every
// kernel that forms candidates calls this, so they agree bitwise.
__device__ __forceinline__ double oc_candidate(uint8_t f, double xe, double ratio, double lam,
                                               double move) {
  if (f != 0) return (f & kVoid) ? 0.0 : 1.0;
  const double r = fmax(ratio / (lam + 1e-30), 1e-30);
  double c = xe * sqrt(r);
  c = fmax(xe - move, fmin(xe + move, c));
  return fmin(fmax(c, 0.0), 1.0);
}

// Stage the stencil in LDS; all threads of the block must call.
__device__ __forceinline__ void stage_stencil(int nsten, const int* __restrict__ off,
                                              const double* __restrict__ w, int* s_off,
                                              double* s_w) {
  for (int i = threadIdx.x; i < nsten; i += blockDim.x) {
    s_off[i] = off[i];
    s_w[i] = w[i];
  }
  __syncthreads();
}

// sum_j w(e, j) v[j] over in-domain neighbours j of element (iy, ix, iz).
__device__ __forceinline__ double stencil_gather(const Grid& g, int iy, int ix, int iz, int nsten,
                                                 const int* s_off, const double* s_w,
                                                 const double* __restrict__ v) {
  double acc = 0.0;
  const int sxy = g.nelx * g.nely;
  for (int k = 0; k < nsten; ++k) {
    int dy, dx, dz;
    decode_offset(s_off[k], dy, dx, dz);
    const int jy = iy + dy, jx = ix + dx, jz = iz + dz;
    if ((unsigned)jy < (unsigned)g.nely && (unsigned)jx < (unsigned)g.nelx &&
        (unsigned)jz < (unsigned)g.nelz)
      acc += s_w[k] * v[jy + jx * g.nely + jz * sxy];
  }
  return acc;
}

// Two fields at once (the transposed chain rule for compliance and volume).
__device__ __forceinline__ void stencil_gather2(const Grid& g, int iy, int ix, int iz, int nsten,
                                                const int* s_off, const double* s_w,
                                                const double* __restrict__ a,
                                                const double* __restrict__ b, double& ra,
                                                double& rb) {
  double sa = 0.0, sb = 0.0;
  const int sxy = g.nelx * g.nely;
  for (int k = 0; k < nsten; ++k) {
    int dy, dx, dz;
    decode_offset(s_off[k], dy, dx, dz);
    const int jy = iy + dy, jx = ix + dx, jz = iz + dz;
    if ((unsigned)jy < (unsigned)g.nely && (unsigned)jx < (unsigned)g.nelx &&
        (unsigned)jz < (unsigned)g.nelz) {
      const int j = jy + jx * g.nely + jz * sxy;
      sa += s_w[k] * a[j];
      sb += s_w[k] * b[j];
    }
  }
  ra = sa;
  rb = sb;
}

// Heaviside projection with protected clamp (pytopo3d/utils/filter.py:53-60,95-98):
//   rho = clip((tanh(b*eta) + tanh(b*(f-eta))) / den, 0, 1),  d = b (1 - t^2) / den,
//   solid -> 1, void -> 0 (void last), d = 0 on both.
__device__ __forceinline__ void project(double f, double beta, double eta, double tb_eta,
                                        double den, uint8_t flag, double& rho, double& d) {
  const double t = tanh(beta * (f - eta));
  double r = (tb_eta + t) / den;
  r = fmin(fmax(r, 0.0), 1.0);
  double dd = beta * (1.0 - t * t) / den;
  if (flag & kSolid) r = 1.0;
  if (flag & kVoid) r = 0.0;
  if (flag & (kSolid | kVoid)) dd = 0.0;
  rho = r;
  d = dd;
}

__device__ __forceinline__ void elem_coords(const Grid& g, int e, int& iy, int& ix, int& iz) {
  iy = e % g.nely;
  const int t = e / g.nely;
  ix = t % g.nelx;
  iz = t / g.nelx;
}

}  // namespace tk
