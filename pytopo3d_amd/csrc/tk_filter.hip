// Density filter, Heaviside projection and the transposed chain rule.
//   forward   : pytopo3d/utils/filter.py:63-100  (build_physical_density)
//   transpose : pytopo3d/utils/filter.py:103-119 (apply_density_filter_chain_rule)
//   row sums  : pytopo3d/utils/filter.py:223-226 (Hs, zero sums become 1)
#include "tk_filter_dev.hpp"
#include "tk_plan.hpp"

namespace tk {

__global__ __launch_bounds__(kBlock) void k_filter_hs(Grid g, int nsten, const int* __restrict__ off,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ Hs) {
  __shared__ int s_off[kMaxStencil];
  __shared__ double s_w[kMaxStencil];
  stage_stencil(nsten, off, w, s_off, s_w);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nele) return;
  int iy, ix, iz;
  elem_coords(g, e, iy, ix, iz);
  double acc = 0.0;
  for (int k = 0; k < nsten; ++k) {
    int dy, dx, dz;
    decode_offset(s_off[k], dy, dx, dz);
    const int jy = iy + dy, jx = ix + dx, jz = iz + dz;
    if ((unsigned)jy < (unsigned)g.nely && (unsigned)jx < (unsigned)g.nelx &&
        (unsigned)jz < (unsigned)g.nelz)
      acc += s_w[k];
  }
  Hs[e] = (acc == 0.0) ? 1.0 : acc;
}

// rho_f = (H x)/Hs, rho = projected + clamped, d = d rho / d rho_f.  Any output may be null.
__global__ __launch_bounds__(kBlock) void k_filter_project(
    Grid g, int nsten, const int* __restrict__ off, const double* __restrict__ w,
    const double* __restrict__ Hs, const uint8_t* __restrict__ fl, const double* __restrict__ x,
    double beta, double eta, double tb_eta, double den, double* __restrict__ rf,
    double* __restrict__ rho, double* __restrict__ drho) {
  __shared__ int s_off[kMaxStencil];
  __shared__ double s_w[kMaxStencil];
  stage_stencil(nsten, off, w, s_off, s_w);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nele) return;
  int iy, ix, iz;
  elem_coords(g, e, iy, ix, iz);
  const double f = stencil_gather(g, iy, ix, iz, nsten, s_off, s_w, x) / Hs[e];
  if (rf) rf[e] = f;
  double r, d;
  project(f, beta, eta, tb_eta, den, fl[e], r, d);
  if (rho) rho[e] = r;
  if (drho) drho[e] = d;
}

// out = H^T v for two fields (H symmetric: H^T v = H v).
__global__ __launch_bounds__(kBlock) void k_filter_transpose2(
    Grid g, int nsten, const int* __restrict__ off, const double* __restrict__ w,
    const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ oa,
    double* __restrict__ ob) {
  __shared__ int s_off[kMaxStencil];
  __shared__ double s_w[kMaxStencil];
  stage_stencil(nsten, off, w, s_off, s_w);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nele) return;
  int iy, ix, iz;
  elem_coords(g, e, iy, ix, iz);
  if (b) {
    double ra, rb;
    stencil_gather2(g, iy, ix, iz, nsten, s_off, s_w, a, b, ra, rb);
    oa[e] = ra;
    ob[e] = rb;
  } else {
    oa[e] = stencil_gather(g, iy, ix, iz, nsten, s_off, s_w, a);
  }
}

// sum over free elements of the projected density of x -> sc->sum{slot}.
__global__ __launch_bounds__(kBlock) void k_free_sum_projected(
    Grid g, int nsten, const int* __restrict__ off, const double* __restrict__ w,
    const double* __restrict__ Hs, const uint8_t* __restrict__ fl, const double* __restrict__ x,
    double beta, double eta, double tb_eta, double den, int slot, double* partials, Scalars* sc) {
  __shared__ int s_off[kMaxStencil];
  __shared__ double s_w[kMaxStencil];
  stage_stencil(nsten, off, w, s_off, s_w);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (e < g.nele && fl[e] == 0) {
    int iy, ix, iz;
    elem_coords(g, e, iy, ix, iz);
    const double f = stencil_gather(g, iy, ix, iz, nsten, s_off, s_w, x) / Hs[e];
    double r, d;
    project(f, beta, eta, tb_eta, den, 0, r, d);
    v[0] = r;
  }
  block_partials<1>(v, partials);
}

// Attainable-volume designs (pytopo3d/core/optimizer.py:501-531):
// lower = 0 on free, 1 on solid, 0 on void; upper = 1 except 0 on void.
__global__ __launch_bounds__(kBlock) void k_bounds_design(int n, const uint8_t* __restrict__ fl,
                                                           int upper, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint8_t f = fl[e];
  double v = upper ? 1.0 : 0.0;
  if (f & kSolid) v = 1.0;
  if (f & kVoid) v = 0.0;
  out[e] = v;
}

__global__ __launch_bounds__(kBlock) void k_count_free(int n, const uint8_t* __restrict__ fl,
                                                        double* partials,
                                                        Scalars* sc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {(e < n && fl[e] == 0) ? 1.0 : 0.0};
  block_partials<1>(v, partials);
}

// ---------------------------------------------------------------------------
void launch_filter_hs(Plan& P, hipStream_t s) {
  hipLaunchKernelGGL(k_filter_hs, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g, P.nsten,
                     P.sten_off, P.sten_w, P.Hs);
}

void launch_filter_project(Plan& P, const double* x, double beta, double eta, double tb_eta,
                           double den, double* rf, double* rho, double* drho, hipStream_t s) {
  P.cnt.filter_passes++;
  if (P.ftiled) {
    tiled_filter_project(P, x, beta, eta, tb_eta, den, rf, rho, drho, s);
    return;
  }
  hipLaunchKernelGGL(k_filter_project, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g,
                     P.nsten, P.sten_off, P.sten_w, P.Hs, P.eflags, x, beta, eta, tb_eta, den,
                     rf, rho, drho);
}

void launch_filter_transpose2(Plan& P, const double* wc, const double* wv, double* oc,
                              double* ov, hipStream_t s) {
  P.cnt.filter_passes++;
  if (P.ftiled) {
    tiled_filter_plain(P, wc, oc, s);
    if (wv) tiled_filter_plain(P, wv, ov, s);
    return;
  }
  hipLaunchKernelGGL(k_filter_transpose2, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g,
                     P.nsten, P.sten_off, P.sten_w, wc, wv, oc, ov);
}

void launch_free_sum_projected(Plan& P, const double* x, double beta, double eta, double tb_eta,
                               double den, int slot, hipStream_t s) {
  P.cnt.filter_passes++;
  if (P.ftiled) {
    tiled_free_sum(P, x, beta, eta, tb_eta, den, slot, s);
    return;
  }
  hipLaunchKernelGGL(k_free_sum_projected, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g,
                     P.nsten, P.sten_off, P.sten_w, P.Hs, P.eflags, x, beta, eta, tb_eta, den,
                     slot, P.partials, P.sc);
  finalize(P, slot == 0 ? kFinSum0 : kFinSum1, nblocks(P.g.nele), s);
}

void launch_bounds_design(Plan& P, int upper, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_design, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     P.eflags, upper, out);
}

void launch_count_free(Plan& P, hipStream_t s) {
  hipLaunchKernelGGL(k_count_free, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     P.eflags, P.partials, P.sc);
  finalize(P, kFinSum0, nblocks(P.g.nele), s);
}

}  // namespace tk
