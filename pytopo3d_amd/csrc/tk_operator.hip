// Matrix-free H8 stiffness operator and the geometric-multigrid transfer,
// smoothing and coarse-solve kernels.
//
// Replaces the assembled K of pytopo3d/core/optimizer.py:584-594 (kron of the
// SIMP scale with KE, np.add.at into CSR, the Kff sub-matrix): K u is applied
// node by node as sum over the <= 8 elements around the node of
// s_e * KE[rows of the node] * u_e, with no atomics.  Fixed DOFs are handled
// as identity rows (inputs masked to 0, outputs y = u), which is the free-DOF
// system of pytopo3d/core/optimizer.py:591-594 padded with the identity.
//
// Multigrid: vertex-centred 2:1 coarsening (ceil for odd counts), trilinear
// prolongation P and restriction R = P^T, coarse element scale
// s_E = (1/4) sum of the child scales (mean of 8 children times 2 because KE
// is proportional to the element size), Chebyshev smoothing with the
// a-priori bound lambda_max(D^-1 K) <= lambda_max(diag(KE)^-1 KE), and a dense
// Cholesky-based inverse on the coarsest grid.
#include "tk_plan.hpp"
#include "tk_transfer.hpp"

#include <type_traits>

namespace tk {

constexpr double kEmin = 1.0e-9;

// ------------------------------------------------------------------ operator core
template <typename T>
__device__ __forceinline__ void stage_ke(const T* __restrict__ KE, T* ke) {
  for (int i = threadIdx.x; i < 576; i += blockDim.x) ke[i] = KE[i];
  __syncthreads();
}

__device__ __forceinline__ void node_coords(const Grid& g, int n, int& iy, int& ix, int& iz) {
  iy = n % g.nny;
  const int t = n / g.nny;
  ix = t % g.nnx;
  iz = t / g.nnx;
}

// acc = (K u)[node (iy, ix, iz)], inputs masked at fixed DOFs.
template <typename T>
__device__ __forceinline__ void node_Ku(const Grid& g, int iy, int ix, int iz, const T* ke,
                                        const T* __restrict__ s, const T* __restrict__ u,
                                        const uint8_t* __restrict__ fixed, T acc[3]) {
  acc[0] = acc[1] = acc[2] = T(0);
  const int sxy = g.nelx * g.nely;
  const int nxy = g.nnx * g.nny;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ey = iy - corner_y(a), ex = ix - corner_x(a), ez = iz - corner_z(a);
    if ((unsigned)ey >= (unsigned)g.nely || (unsigned)ex >= (unsigned)g.nelx ||
        (unsigned)ez >= (unsigned)g.nelz)
      continue;
    const T se = s[ey + ex * g.nely + ez * sxy];
    T p0 = T(0), p1 = T(0), p2 = T(0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int m = (ey + corner_y(b)) + (ex + corner_x(b)) * g.nny + (ez + corner_z(b)) * nxy;
      const uint8_t fm = fixed[m];
      const T u0 = (fm & 1) ? T(0) : u[3 * m];
      const T u1 = (fm & 2) ? T(0) : u[3 * m + 1];
      const T u2 = (fm & 4) ? T(0) : u[3 * m + 2];
      const T* r0 = ke + (3 * a) * 24 + 3 * b;
      p0 += r0[0] * u0 + r0[1] * u1 + r0[2] * u2;
      p1 += r0[24] * u0 + r0[25] * u1 + r0[26] * u2;
      p2 += r0[48] * u0 + r0[49] * u1 + r0[50] * u2;
    }
    acc[0] += se * p0;
    acc[1] += se * p1;
    acc[2] += se * p2;
  }
}

// y = K u (identity on fixed DOFs); optional partial u.y -> CG step length.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_apply(Grid g, const T* __restrict__ KE,
                                                   const T* __restrict__ s,
                                                   const uint8_t* __restrict__ fixed,
                                                   const T* __restrict__ u, T* __restrict__ y,
                                                   int with_dot, const int* __restrict__ done,
                                                   double* partials,
                                                   Scalars* sc) {
  if (done && *done) return;
  __shared__ T ke[576];
  stage_ke(KE, ke);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double dot[1] = {0.0};
  if (n < g.nnode) {
    int iy, ix, iz;
    node_coords(g, n, iy, ix, iz);
    T acc[3];
    node_Ku(g, iy, ix, iz, ke, s, u, fixed, acc);
    const uint8_t f = fixed[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T uc = u[3 * n + c];
      const T yc = ((f >> c) & 1) ? uc : acc[c];
      y[3 * n + c] = yc;
      dot[0] += (double)uc * (double)yc;
    }
  }
  if (with_dot) block_partials<1>(dot, partials);
}

// r = b - K x on free DOFs, r = 0 on fixed DOFs (the restriction relies on it)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(Grid g, const T* __restrict__ KE,
                                                      const T* __restrict__ s,
                                                      const uint8_t* __restrict__ fixed,
                                                      const T* __restrict__ b,
                                                      const T* __restrict__ x, T* __restrict__ r,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ T ke[576];
  stage_ke(KE, ke);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  int iy, ix, iz;
  node_coords(g, n, iy, ix, iz);
  T acc[3];
  node_Ku(g, iy, ix, iz, ke, s, x, fixed, acc);
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[3 * n + c] = ((f >> c) & 1) ? T(0) : b[3 * n + c] - acc[c];
}

// First Chebyshev step: d = c0 Dinv (b - K x); xo = x + d  (x = 0: d = c0 Dinv b).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_first(Grid g, const T* __restrict__ KE,
                                                        const T* __restrict__ s,
                                                        const uint8_t* __restrict__ fixed,
                                                        const T* __restrict__ dinv,
                                                        const T* __restrict__ b,
                                                        const T* __restrict__ x,
                                                        T* __restrict__ xo, T* __restrict__ d,
                                                        T c0, int x_zero,
                                                        const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ T ke[576];
  if (!x_zero) stage_ke(KE, ke);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  if (x_zero) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int i = 3 * n + c;
      const T dd = c0 * dinv[i] * b[i];
      d[i] = dd;
      xo[i] = dd;
    }
    return;
  }
  int iy, ix, iz;
  node_coords(g, n, iy, ix, iz);
  T acc[3];
  node_Ku(g, iy, ix, iz, ke, s, x, fixed, acc);
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int i = 3 * n + c;
    const T yc = ((f >> c) & 1) ? x[i] : acc[c];
    const T dd = c0 * dinv[i] * (b[i] - yc);
    d[i] = dd;
    xo[i] = x[i] + dd;
  }
}

// Chebyshev step k >= 1: d = c1 d + c2 Dinv (b - K x); xo = x + d.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_step(Grid g, const T* __restrict__ KE,
                                                       const T* __restrict__ s,
                                                       const uint8_t* __restrict__ fixed,
                                                       const T* __restrict__ dinv,
                                                       const T* __restrict__ b,
                                                       const T* __restrict__ x,
                                                       T* __restrict__ xo, T* __restrict__ d,
                                                       T c1, T c2, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ T ke[576];
  stage_ke(KE, ke);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  int iy, ix, iz;
  node_coords(g, n, iy, ix, iz);
  T acc[3];
  node_Ku(g, iy, ix, iz, ke, s, x, fixed, acc);
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int i = 3 * n + c;
    const T yc = ((f >> c) & 1) ? x[i] : acc[c];
    const T dd = c1 * d[i] + c2 * dinv[i] * (b[i] - yc);
    d[i] = dd;
    xo[i] = x[i] + dd;
  }
}

// inverse diagonal: diag = sum_e s_e KE[ii]; 1 on fixed DOFs.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_inv(Grid g, const double* __restrict__ KE,
                                                      const T* __restrict__ s,
                                                      const uint8_t* __restrict__ fixed,
                                                      T* __restrict__ dinv,
                                                      T* __restrict__ diag_out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  int iy, ix, iz;
  node_coords(g, n, iy, ix, iz);
  const int sxy = g.nelx * g.nely;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ey = iy - corner_y(a), ex = ix - corner_x(a), ez = iz - corner_z(a);
    if ((unsigned)ey >= (unsigned)g.nely || (unsigned)ex >= (unsigned)g.nelx ||
        (unsigned)ez >= (unsigned)g.nelz)
      continue;
    const double se = (double)s[ey + ex * g.nely + ez * sxy];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += se * KE[(3 * a + c) * 25];
  }
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const bool fx = (f >> c) & 1;
    const double dg = fx ? 1.0 : acc[c];
    if (dinv) dinv[3 * n + c] = (T)(1.0 / dg);
    if (diag_out) diag_out[3 * n + c] = (T)dg;
  }
}

// k_diag_inv of every level in one launch: the blocks of level l are
// [block0[l], block0[l + 1]).
template <typename T>
struct DiagLevels {
  int nlev;
  int block0[kMaxLevels + 1];
  Grid g[kMaxLevels];
  const T* s[kMaxLevels];
  const uint8_t* fixed[kMaxLevels];
  T* dinv[kMaxLevels];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_inv_levels(const double* __restrict__ KE,
                                                             DiagLevels<T> d) {
  int l = 0;
  while (l + 1 < d.nlev && (int)blockIdx.x >= d.block0[l + 1]) ++l;
  const Grid g = d.g[l];
  const int n = ((int)blockIdx.x - d.block0[l]) * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  const T* __restrict__ s = d.s[l];
  int iy, ix, iz;
  node_coords(g, n, iy, ix, iz);
  const int sxy = g.nelx * g.nely;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ey = iy - corner_y(a), ex = ix - corner_x(a), ez = iz - corner_z(a);
    if ((unsigned)ey >= (unsigned)g.nely || (unsigned)ex >= (unsigned)g.nelx ||
        (unsigned)ez >= (unsigned)g.nelz)
      continue;
    const double se = (double)s[ey + ex * g.nely + ez * sxy];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += se * KE[(3 * a + c) * 25];
  }
  const uint8_t f = d.fixed[l][n];
  T* __restrict__ dinv = d.dinv[l];
#pragma unroll
  for (int c = 0; c < 3; ++c) dinv[3 * n + c] = (T)(1.0 / (((f >> c) & 1) ? 1.0 : acc[c]));
}

// s = Emin + (1 - Emin) rho^p  (pytopo3d/core/optimizer.py:584-586)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_stiffness_scale(int n, const double* __restrict__ rho,
                                                             double penal, T* __restrict__ s) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  s[e] = (T)(kEmin + (1.0 - kEmin) * pow(rho[e], penal));
}

// The same in FP64 and its FP32 rounding in one pass.
__global__ __launch_bounds__(kBlock) void k_stiffness_scale2(int n, const double* __restrict__ rho,
                                                              double penal, double* __restrict__ s64,
                                                              float* __restrict__ s32) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double v = kEmin + (1.0 - kEmin) * pow(rho[e], penal);
  s64[e] = v;
  s32[e] = (float)v;
}

// s_c[E] = 1/4 * sum of the (in-domain) children scales.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_coarsen_scale(Grid f, Grid c, const T* __restrict__ sf,
                                                           T* __restrict__ sc) {
  const int E = blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= c.nele) return;
  const int Ey = E % c.nely, t = E / c.nely, Ex = t % c.nelx, Ez = t / c.nelx;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ey = 2 * Ey + (k & 1), ex = 2 * Ex + ((k >> 1) & 1), ez = 2 * Ez + (k >> 2);
    if (ey < f.nely && ex < f.nelx && ez < f.nelz)
      acc += (double)sf[ey + ex * f.nely + ez * f.nelx * f.nely];
  }
  sc[E] = (T)(0.25 * acc);
}

// k_coarsen_scale for the FP64 and the FP32 chain in one pass (each from its
// own finer level, so both match the separate passes bit for bit).
__global__ __launch_bounds__(kBlock) void k_coarsen_scale2(Grid f, Grid c,
                                                            const double* __restrict__ sf64,
                                                            const float* __restrict__ sf32,
                                                            double* __restrict__ sc64,
                                                            float* __restrict__ sc32) {
  const int E = blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= c.nele) return;
  const int Ey = E % c.nely, t = E / c.nely, Ex = t % c.nelx, Ez = t / c.nelx;
  double a64 = 0.0, a32 = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ey = 2 * Ey + (k & 1), ex = 2 * Ex + ((k >> 1) & 1), ez = 2 * Ez + (k >> 2);
    if (ey < f.nely && ex < f.nelx && ez < f.nelz) {
      const int e = ey + ex * f.nely + ez * f.nelx * f.nely;
      a64 += sf64[e];
      a32 += (double)sf32[e];
    }
  }
  sc64[E] = 0.25 * a64;
  sc32[E] = (float)(0.25 * a32);
}

__global__ __launch_bounds__(kBlock) void k_coarsen_fixed(Grid f, Grid c,
                                                           const uint8_t* __restrict__ ff,
                                                           uint8_t* __restrict__ fc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= c.nnode) return;
  int Iy, Ix, Iz;
  node_coords(c, I, Iy, Ix, Iz);
  const int iy = 2 * Iy, ix = 2 * Ix, iz = 2 * Iz;
  uint8_t v = 0;
  if (iy < f.nny && ix < f.nnx && iz < f.nnz) v = ff[iy + ix * f.nny + iz * f.nnx * f.nny];
  fc[I] = v;
}

// b_c = R r_f = P^T r_f (trilinear weights), zero on coarse fixed DOFs.  The
// fine residual is zero on fixed fine DOFs (the residual kernels write 0
// there), so no fine mask is read.  Each coarse node gathers its 3x3x3 fine
// neighbourhood with whole-node (3-component) loads; neighbouring threads
// share fine nodes through L1.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(Grid f, Grid c,
                                                      const uint8_t* __restrict__ fc,
                                                      const T* __restrict__ rf,
                                                      T* __restrict__ bc,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  const int I = xcd_logical(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (I >= c.nnode) return;
  restrict_node(f, c, fc, rf, bc, I);
}

// x_f += P x_c at the fine nodes (2Y, ix, iz) and (2Y + 1, ix, iz): both
// interpolate between coarse y rows Y and Y + 1 (the even node with weights 1
// and 0), so the pair shares one set of 8 coarse gathers, interpolated in x
// and z first.  Free DOFs only.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolong_add2(Grid c, Grid f,
                                                          const uint8_t* __restrict__ ff,
                                                          const T* __restrict__ xc,
                                                          T* __restrict__ xf,
                                                          const int* __restrict__ done) {
  if (done && *done) return;
  const int npy = (f.nny + 1) >> 1;
  const int i = xcd_logical(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= npy * f.nnx * f.nnz) return;
  const int Y = i % npy, t = i / npy, ix = t % f.nnx, iz = t / f.nnx;
  const int Y1 = Y + 1 < c.nny ? Y + 1 : Y;
  const int x0 = ix >> 1, x1 = (ix + 1) >> 1, z0 = iz >> 1, z1 = (iz + 1) >> 1;
  const T wx[2] = {(ix & 1) ? T(0.5) : T(1), (ix & 1) ? T(0.5) : T(0)};
  const T wz[2] = {(iz & 1) ? T(0.5) : T(1), (iz & 1) ? T(0.5) : T(0)};
  const int cxy = c.nnx * c.nny;
  T a0[3] = {T(0), T(0), T(0)}, a1[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int kx = k & 1, kz = k >> 1;
    const int base = (kx ? x1 : x0) * c.nny + (kz ? z1 : z0) * cxy;
    const T w = wx[kx] * wz[kz];
    const T* q0 = xc + 3 * (size_t)(Y + base);
    const T* q1 = xc + 3 * (size_t)(Y1 + base);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a0[r] += w * q0[r];
      a1[r] += w * q1[r];
    }
  }
  const int n0 = 2 * Y + ix * f.nny + iz * f.nnx * f.nny;
  // both nodes' values loaded before either is stored (unconditional loads;
  // the odd node, when the grid has none, at the even one's clamped index)
  const bool odd = 2 * Y + 1 < f.nny;
  const int n1 = odd ? n0 + 1 : n0;
  T* o0 = xf + 3 * (size_t)n0;
  T* o1 = xf + 3 * (size_t)n1;
  const uint8_t f0 = ff[n0], f1 = ff[n1];
  T v0[3], v1[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    v0[r] = o0[r];
    v1[r] = o1[r];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) o0[r] = ((f0 >> r) & 1) ? v0[r] : v0[r] + a0[r];
  if (odd)
#pragma unroll
    for (int r = 0; r < 3; ++r) o1[r] = ((f1 >> r) & 1) ? v1[r] : v1[r] + T(0.5) * (a0[r] + a1[r]);
}

// ------------------------------------------------------------------ coarsest grid
__device__ __forceinline__ int pk(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

__device__ __forceinline__ int coarse_dof(const Grid& g, int E, int a, int c) {
  const int Ey = E % g.nely, t = E / g.nely, Ex = t % g.nelx, Ez = t / g.nelx;
  const int m = (Ey + corner_y(a)) + (Ex + corner_x(a)) * g.nny + (Ez + corner_z(a)) * g.nnx * g.nny;
  return 3 * m + c;
}

// One workgroup: assemble the dense coarsest operator in LDS (packed lower
// triangle), impose identity rows on fixed DOFs, factorise (Cholesky) and
// write the dense inverse to global memory.
__global__ __launch_bounds__(1024) void k_coarse_setup(Grid g, const double* __restrict__ KE,
                                                        const double* __restrict__ s,
                                                        const uint8_t* __restrict__ fixed,
                                                        double* __restrict__ inv,
                                                        double* __restrict__ Lg) {
  extern __shared__ double A[];
  const int n = g.ndof;
  const int np = n * (n + 1) / 2;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < np; i += nt) A[i] = 0.0;
  __syncthreads();
  for (int E = 0; E < g.nele; ++E) {
    const double se = s[E];
    for (int t = tid; t < 576; t += nt) {
      const int i = t / 24, j = t % 24;
      const int gi = coarse_dof(g, E, i / 3, i % 3), gj = coarse_dof(g, E, j / 3, j % 3);
      if (gi >= gj) A[pk(gi, gj)] += se * KE[t];
    }
    __syncthreads();
  }
  for (int t = tid; t < n * n; t += nt) {
    const int i = t / n, j = t % n;
    if (j > i) continue;
    const bool fi = (fixed[i / 3] >> (i % 3)) & 1, fj = (fixed[j / 3] >> (j % 3)) & 1;
    if (fi || fj) A[pk(i, j)] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  // reference scale for pivot regularisation
  __shared__ double dmax;
  if (tid == 0) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = fmax(m, A[pk(i, i)]);
    dmax = m;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      double p = A[pk(k, k)];
      if (!(p > 1e-13 * dmax)) p = 1e-13 * dmax + 1e-300;  // singular coarse space guard
      A[pk(k, k)] = sqrt(p);
    }
    __syncthreads();
    const double lkk = A[pk(k, k)];
    for (int i = k + 1 + tid; i < n; i += nt) A[pk(i, k)] /= lkk;
    __syncthreads();
    const int m = n - k - 1;
    for (int t = tid; t < m * m; t += nt) {
      const int ii = t / m, jj = t % m;
      if (jj > ii) continue;
      const int i = k + 1 + ii, j = k + 1 + jj;
      A[pk(i, j)] -= A[pk(i, k)] * A[pk(j, k)];
    }
    __syncthreads();
  }
  for (int t = tid; t < np; t += nt) Lg[t] = A[t];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Column j of A^-1 = L^-T L^-1 e_j, one wave per column: forward and backward
// substitution with the inner products spread over the 64 lanes.  A^-1 is
// symmetric, so column j is stored as row j (coalesced by the coarse solve).
__global__ __launch_bounds__(64) void k_coarse_inverse(int n, const double* __restrict__ Lg,
                                                        double* __restrict__ inv64,
                                                        float* __restrict__ inv32) {
  __shared__ double y[kCoarseMaxDof];
  const int j = blockIdx.x, lane = threadIdx.x;
  for (int i = lane; i < n; i += 64) y[i] = 0.0;
  __syncthreads();
  for (int i = j; i < n; ++i) {
    double acc = 0.0;
    for (int k = j + lane; k < i; k += 64) acc += Lg[pk(i, k)] * y[k];
    acc = wave_sum(acc);
    if (lane == 0) y[i] = ((i == j ? 1.0 : 0.0) - acc) / Lg[pk(i, i)];
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = 0.0;
    for (int k = i + 1 + lane; k < n; k += 64) acc += Lg[pk(k, i)] * y[k];
    acc = wave_sum(acc);
    if (lane == 0) y[i] = (y[i] - acc) / Lg[pk(i, i)];
    __syncthreads();
  }
  for (int i = lane; i < n; i += 64) {
    inv64[j * n + i] = y[i];
    inv32[j * n + i] = (float)y[i];
  }
}

// x = A^-1 b on the coarsest grid: one wave per row.
template <typename T>
__global__ __launch_bounds__(64) void k_coarse_solve(int n, const T* __restrict__ inv,
                                                      const T* __restrict__ b, T* __restrict__ x,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += (double)inv[i * n + j] * (double)b[j];
  acc = wave_sum(acc);
  if (lane == 0) x[i] = (T)acc;
}

// ------------------------------------------------------------------ launchers
template <typename T> static const T* ke_of(const Plan& P);
template <> const double* ke_of<double>(const Plan& P) { return P.KE64; }
template <> const float* ke_of<float>(const Plan& P) { return P.KE32; }

static const int* done_ptr(Plan& P) { return &P.sc->done; }

template <typename T>
void launch_stiffness_scale(Plan& P, const double* rho, double penal, T* s_out, hipStream_t s) {
  hipLaunchKernelGGL(k_stiffness_scale<T>, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     rho, penal, s_out);
}

void launch_stiffness_scale2(Plan& P, const double* rho, double penal, double* s64, float* s32,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_stiffness_scale2, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     rho, penal, s64, s32);
}

template <typename T>
void launch_coarsen_scale(const Level& f, const Level& c, const T* sf, T* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_coarsen_scale<T>, dim3(nblocks(c.g.nele)), dim3(kBlock), 0, s, f.g, c.g, sf,
                     sc);
}

void launch_coarsen_scale2(Plan& P, hipStream_t s) {
  for (int l = 0; l + 1 < P.nlev; ++l) {
    const Level& f = P.lv[l];
    const Level& c = P.lv[l + 1];
    hipLaunchKernelGGL(k_coarsen_scale2, dim3(nblocks(c.g.nele)), dim3(kBlock), 0, s, f.g, c.g,
                       f.f64.s, f.f32.s, c.f64.s, c.f32.s);
  }
}

template <typename T>
void launch_diag_inv_levels(Plan& P, hipStream_t s) {
  DiagLevels<T> d{};
  d.nlev = P.nlev;
  int nb = 0;
  for (int l = 0; l < P.nlev; ++l) {
    const Level& L = P.lv[l];
    const LevelBuf<T>* B;
    if constexpr (std::is_same_v<T, double>) B = &L.f64;
    else B = &L.f32;
    d.block0[l] = nb;
    d.g[l] = L.g;
    d.s[l] = B->s;
    d.fixed[l] = L.fixed;
    d.dinv[l] = B->dinv;
    nb += nblocks(L.g.nnode);
  }
  d.block0[P.nlev] = nb;
  hipLaunchKernelGGL(k_diag_inv_levels<T>, dim3(nb), dim3(kBlock), 0, s, P.KE64, d);
}
template void launch_diag_inv_levels<double>(Plan&, hipStream_t);
template void launch_diag_inv_levels<float>(Plan&, hipStream_t);

void launch_coarsen_fixed(const Level& f, const Level& c, hipStream_t s) {
  hipLaunchKernelGGL(k_coarsen_fixed, dim3(nblocks(c.g.nnode)), dim3(kBlock), 0, s, f.g, c.g,
                     f.fixed, c.fixed);
}

template <typename T>
void launch_diag_inv(Plan& P, const Level& L, const T* sl, T* dinv, T* diag_out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_inv<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g, P.KE64, sl,
                     L.fixed, dinv, diag_out);
}

template <typename T>
void launch_apply(Plan& P, const Level& L, const T* sl, const T* u, T* y, double* /*unused*/,
                  int with_dot, hipStream_t s) {
  if (&L == &P.lv[0]) {
    if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
  }
  if (P.use_el) {
    el_apply<T>(P, L, sl, u, y, with_dot, with_dot ? done_ptr(P) : nullptr, s);
    return;
  }
  hipLaunchKernelGGL(k_apply<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g, ke_of<T>(P), sl,
                     L.fixed, u, y, with_dot, with_dot ? done_ptr(P) : nullptr, P.partials, P.sc);
  if (with_dot) finalize(P, kFinAlpha, nblocks(L.g.nnode), s);
}

template <typename T>
void launch_residual(Plan& P, const Level& L, const T* sl, const T* b, const T* x, T* r,
                     hipStream_t s, const int* done) {
  if (&L == &P.lv[0]) {
    if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
  }
  if (P.use_el) {
    el_residual<T>(P, L, sl, b, x, r, done, s);
    return;
  }
  hipLaunchKernelGGL(k_residual<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g, ke_of<T>(P),
                     sl, L.fixed, b, x, r, done);
}

template <typename T>
void launch_cheb_first(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, const T* x,
                       T* xo, T* d, double c0, int x_zero, int write_d, hipStream_t s) {
  if (&L == &P.lv[0] && !x_zero) {
    if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
  }
  if (P.use_el && !x_zero) {
    el_cheb<T>(P, L, sl, dinv, b, x, xo, d, c0, 0.0, 0.0, 1, write_d, done_ptr(P), s);
    return;
  }
  hipLaunchKernelGGL(k_cheb_first<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g,
                     ke_of<T>(P), sl, L.fixed, dinv, b, x, xo, d, (T)c0, x_zero, done_ptr(P));
}

template <typename T>
void launch_cheb_step(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, const T* x,
                      T* xo, T* d, double c1, double c2, int write_d, hipStream_t s,
                      int* dot_blocks) {
  if (&L == &P.lv[0]) {
    if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
  }
  if (P.use_el) {
    el_cheb<T>(P, L, sl, dinv, b, x, xo, d, 0.0, c1, c2, 0, write_d, done_ptr(P), s, dot_blocks);
    return;
  }
  if (dot_blocks) *dot_blocks = 0;   // the node kernel forms no partials
  hipLaunchKernelGGL(k_cheb_step<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g,
                     ke_of<T>(P), sl, L.fixed, dinv, b, x, xo, d, (T)c1, (T)c2, done_ptr(P));
}

template <typename T>
void launch_restrict(const Level& f, const Level& c, const T* rf, T* bc, hipStream_t s,
                     const int* done) {
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblocks(c.g.nnode)), dim3(kBlock), 0, s, f.g, c.g,
                     c.fixed, rf, bc, done);
}

template <typename T>
void launch_prolong_add(const Level& c, const Level& f, const T* xc, T* xf, hipStream_t s,
                        const int* done) {
  const int pairs = (f.g.nny + 1) / 2 * f.g.nnx * f.g.nnz;
  hipLaunchKernelGGL(k_prolong_add2<T>, dim3(nblocks(pairs)), dim3(kBlock), 0, s, c.g, f.g,
                     f.fixed, xc, xf, done);
}

void launch_coarse_setup(Plan& P, hipStream_t s) {
  const Level& C = P.lv[P.nlev - 1];
  const int n = C.g.ndof;
  const size_t lds = sizeof(double) * (size_t)n * (n + 1) / 2;
  hipLaunchKernelGGL(k_coarse_setup, dim3(1), dim3(1024), lds, s, C.g, P.KE64, C.f64.s, C.fixed,
                     P.coarse_inv64, P.coarse_L);
  hipLaunchKernelGGL(k_coarse_inverse, dim3(n), dim3(64), 0, s, n, P.coarse_L, P.coarse_inv64,
                     P.coarse_inv32);
}

template <typename T> static const T* inv_of(const Plan& P);
template <> const double* inv_of<double>(const Plan& P) { return P.coarse_inv64; }
template <> const float* inv_of<float>(const Plan& P) { return P.coarse_inv32; }

template <typename T>
void launch_coarse_solve(Plan& P, const T* b, T* x, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse_solve<T>, dim3(P.ncoarse), dim3(64), 0, s, P.ncoarse,
                     inv_of<T>(P), b, x, done_ptr(P));
}

hipError_t coarse_setup_attr(size_t lds) {
  return hipFuncSetAttribute((const void*)k_coarse_setup,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// explicit instantiations
#define TK_INST(T)                                                                                  \
  template void launch_stiffness_scale<T>(Plan&, const double*, double, T*, hipStream_t);          \
  template void launch_coarsen_scale<T>(const Level&, const Level&, const T*, T*, hipStream_t);    \
  template void launch_diag_inv<T>(Plan&, const Level&, const T*, T*, T*, hipStream_t);            \
  template void launch_apply<T>(Plan&, const Level&, const T*, const T*, T*, double*, int,         \
                                hipStream_t);                                                      \
  template void launch_residual<T>(Plan&, const Level&, const T*, const T*, const T*, T*,          \
                                   hipStream_t, const int*);                                       \
  template void launch_cheb_first<T>(Plan&, const Level&, const T*, const T*, const T*, const T*,  \
                                     T*, T*, double, int, int, hipStream_t);                       \
  template void launch_cheb_step<T>(Plan&, const Level&, const T*, const T*, const T*, const T*,   \
                                    T*, T*, double, double, int, hipStream_t, int*);               \
  template void launch_restrict<T>(const Level&, const Level&, const T*, T*, hipStream_t,          \
                                   const int*);                                                    \
  template void launch_prolong_add<T>(const Level&, const Level&, const T*, T*, hipStream_t,       \
                                      const int*);                                                 \
  template void launch_coarse_solve<T>(Plan&, const T*, T*, hipStream_t);
TK_INST(double)
TK_INST(float)
#undef TK_INST

}  // namespace tk
