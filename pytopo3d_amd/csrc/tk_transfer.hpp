// Grid-transfer helpers shared by the transfer kernels and the fused
// smoothing sweeps.
#pragma once

#include "tk_common.hpp"

namespace tk {

__device__ __forceinline__ void grid_node_coords(const Grid& g, int n, int& iy, int& ix, int& iz) {
  iy = n % g.nny;
  const int t = n / g.nny;
  ix = t % g.nnx;
  iz = t / g.nnx;
}

// (R rf) at coarse node (Iy, Ix, Iz): trilinear restriction (R = P^T) of a
// fine field that is zero on fixed fine DOFs, zero on the coarse node's fixed
// DOFs (fixed bits fI).
template <typename T>
__device__ __forceinline__ void restrict_at(const Grid& f, int Iy, int Ix, int Iz, uint8_t fI,
                                            const T* rf, T out[3]) {
  // fine neighbours 2I - 1, 2I, 2I + 1 per axis; indices outside the fine grid
  // are clamped and weighted 0, so all 27 loads are unconditional
  int jy[3], jx[3], jz[3];
  T wy[3], wx[3], wz[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T w = (k == 1) ? T(1) : T(0.5);
    int t = 2 * Iy + k - 1;
    wy[k] = ((unsigned)t < (unsigned)f.nny) ? w : T(0);
    jy[k] = min(max(t, 0), f.nny - 1);
    t = 2 * Ix + k - 1;
    wx[k] = ((unsigned)t < (unsigned)f.nnx) ? w : T(0);
    jx[k] = min(max(t, 0), f.nnx - 1);
    t = 2 * Iz + k - 1;
    wz[k] = ((unsigned)t < (unsigned)f.nnz) ? w : T(0);
    jz[k] = min(max(t, 0), f.nnz - 1);
  }
  T a0 = T(0), a1 = T(0), a2 = T(0);
  const int nxy = f.nnx * f.nny;
#pragma unroll
  for (int kz = 0; kz < 3; ++kz)
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const T wzx = wz[kz] * wx[kx];
      const T* row = rf + 3 * ((size_t)jx[kx] * f.nny + (size_t)jz[kz] * nxy);
#pragma unroll
      for (int ky = 0; ky < 3; ++ky) {
        const T wgt = wzx * wy[ky];
        const T* q = row + 3 * jy[ky];
        a0 += wgt * q[0];
        a1 += wgt * q[1];
        a2 += wgt * q[2];
      }
    }
  out[0] = (fI & 1) ? T(0) : a0;
  out[1] = (fI & 2) ? T(0) : a1;
  out[2] = (fI & 4) ? T(0) : a2;
}

template <typename T>
__device__ __forceinline__ void restrict_node(const Grid& f, const Grid& c,
                                              const uint8_t* __restrict__ fc, const T* rf, T* bc,
                                              int I) {
  int Iy, Ix, Iz;
  grid_node_coords(c, I, Iy, Ix, Iz);
  T o[3];
  restrict_at(f, Iy, Ix, Iz, fc[I], rf, o);
  T* out = bc + 3 * (size_t)I;
  out[0] = o[0];
  out[1] = o[1];
  out[2] = o[2];
}

}  // namespace tk
