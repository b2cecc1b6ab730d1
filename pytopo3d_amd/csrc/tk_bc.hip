// Boundary conditions built on the device from the reference's own inputs:
//   F          build_force_vector   (pytopo3d/utils/assembly.py:78-184)
//   node flags build_supports       (pytopo3d/utils/assembly.py:224-348)
//   elem flags protected masks      (pytopo3d/core/optimizer.py:343-405)
// The reference builds F and the supports with Python loops over every
// element; here one thread per node gathers from its (at most 8) elements and
// one thread per element forms its flags.  Input masks and the force field
// keep the reference's array layout: C-contiguous (nely, nelx, nelz) bytes and
// (nely, nelx, nelz, 3) doubles.
#include "tk_plan.hpp"

namespace tk {

__device__ __forceinline__ int c_index(const Grid& g, int iy, int ix, int iz) {
  return (iy * g.nelx + ix) * g.nelz + iz;   // C order of (nely, nelx, nelz)
}

// partials: count of support-masked elements (an all-False mask means the
// default x = 0 face, pytopo3d/utils/assembly.py:291-303)
__global__ __launch_bounds__(kBlock) void k_bc_count(int n, const uint8_t* __restrict__ m,
                                                     double* partials) {
  double v[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[0] += m[i] ? 1.0 : 0.0;
  block_partials<1>(v, partials);
}

// One thread per node: F from the element force field (or the default line
// load) and the fixed-DOF bits.
__global__ __launch_bounds__(kBlock) void k_bc_nodes(Grid g, const double* __restrict__ ff,
                                                     const uint8_t* __restrict__ sm,
                                                     const Scalars* __restrict__ sc,
                                                     double* __restrict__ F,
                                                     uint8_t* __restrict__ fixed) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  const int iy = n % g.nny, t = n / g.nny, ix = t % g.nnx, iz = t / g.nnx;
  if (F) {
    double f[3] = {0.0, 0.0, 0.0};
    if (ff) {
      // the reference visits elements z, x, y ascending and adds each
      // non-zero element's force / 8 to its corners: the same additions in
      // the same order per node
      for (int ez = iz - 1; ez <= iz; ++ez) {
        if (ez < 0 || ez >= g.nelz) continue;
        for (int ex = ix - 1; ex <= ix; ++ex) {
          if (ex < 0 || ex >= g.nelx) continue;
          for (int ey = iy - 1; ey <= iy; ++ey) {
            if (ey < 0 || ey >= g.nely) continue;
            const double* e = ff + 3LL * c_index(g, ey, ex, ez);
            const double a = e[0], b = e[1], c = e[2];
            if (a != 0.0 || b != 0.0 || c != 0.0) {   // np.any (NaN counts as non-zero)
              f[0] += a / 8.0;
              f[1] += b / 8.0;
              f[2] += c / 8.0;
            }
          }
        }
      }
    } else if (ix == g.nelx && iz == 0) {
      f[2] = -1.0;   // default: -1 on the z DOF of the nodes of x = nelx, z = 0
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) F[3 * n + c] = f[c];
  }
  if (fixed) {
    bool fx = false;
    if (sm && sc->sum0 > 0.0) {
      for (int ez = iz - 1; ez <= iz && !fx; ++ez) {
        if (ez < 0 || ez >= g.nelz) continue;
        for (int ex = ix - 1; ex <= ix; ++ex) {
          if (ex < 0 || ex >= g.nelx) continue;
          for (int ey = iy - 1; ey <= iy; ++ey) {
            if (ey < 0 || ey >= g.nely) continue;
            fx = fx || sm[c_index(g, ey, ex, ez)] != 0;
          }
        }
      }
    } else {
      fx = ix == 0;   // default (or all-False mask): every DOF on the x = 0 face
    }
    fixed[n] = fx ? 7 : 0;
  }
}

// One thread per element (element order out): protected solid = protected
// zone | loaded elements (or the default load strip x = nelx - 1, z = 0) |
// supported elements (or the default strip x = 0); void = obstacle, and void
// wins.
__global__ __launch_bounds__(kBlock) void k_bc_elems(Grid g, const double* __restrict__ ff,
                                                     const uint8_t* __restrict__ sm,
                                                     const uint8_t* __restrict__ ob,
                                                     const uint8_t* __restrict__ pz,
                                                     const Scalars* __restrict__ sc,
                                                     uint8_t* __restrict__ flags) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nele) return;
  const int iy = e % g.nely, t = e / g.nely, ix = t % g.nelx, iz = t / g.nelx;
  const int ci = c_index(g, iy, ix, iz);
  bool solid = pz && pz[ci] != 0;
  if (ff) {
    const double* f = ff + 3LL * ci;
    solid = solid || f[0] != 0.0 || f[1] != 0.0 || f[2] != 0.0;
  } else {
    solid = solid || (ix == g.nelx - 1 && iz == 0);
  }
  if (sm && sc->sum0 > 0.0) solid = solid || sm[ci] != 0;
  else solid = solid || ix == 0;
  const bool void_ = ob && ob[ci] != 0;
  flags[e] = void_ ? kVoid : (solid ? kSolid : 0);
}

void launch_boundary_conditions(Plan& P, const double* ff, const uint8_t* sm, const uint8_t* ob,
                                const uint8_t* pz, double* F, uint8_t* fixed, uint8_t* flags,
                                hipStream_t s) {
  const Grid& g = P.g;
  if (sm) {
    const int G = red_blocks(g.nele);
    hipLaunchKernelGGL(k_bc_count, dim3(G), dim3(kBlock), 0, s, g.nele, sm, P.partials);
    finalize(P, kFinSum0, G, s);
  }
  if (F || fixed)
    hipLaunchKernelGGL(k_bc_nodes, dim3(nblocks(g.nnode)), dim3(kBlock), 0, s, g, ff, sm, P.sc, F,
                       fixed);
  if (flags)
    hipLaunchKernelGGL(k_bc_elems, dim3(nblocks(g.nele)), dim3(kBlock), 0, s, g, ff, sm, ob, pz,
                       P.sc, flags);
}

}  // namespace tk
