// Projected-volume optimality-criteria update, run entirely on the device
// (pytopo3d/utils/oc_update.py:25-124):
//   ratio = max(-dc / max(dv, 1e-20), 0) over free elements, normalised by its
//   maximum (all ones when the maximum is 0);
//   lambda bisection in [0, 1e12]: candidate = clip(max(x - m, min(x + m,
//   x sqrt(max(ratio / (lambda + 1e-30), 1e-30)))), 0, 1) on free elements,
//   1 / 0 on protected solid / void; V = sum over free of the projected
//   filtered candidate; the lambda with the smallest |V - target| (strict <)
//   is kept; V > target raises the lower bound; stop when
//   (hi - lo) / max(hi + lo, tiny) < tol or after max_steps.
// Each evaluation is a (candidate, projected volume, finalize) triple; the
// finalize records V and places the next evaluation (tk_bisect.hpp: midpoints
// whose decision an earlier evaluation implies are not measured, and a few
// probes near the crossing make most of them so).  The kept candidate is
// rebuilt from the kept lambda with the same formula (oc_candidate), so it is
// bitwise the candidate that was measured.
// Also: change / volume / gray statistics (pytopo3d/core/optimizer.py:642-655).
#include "tk_bisect.hpp"
#include "tk_filter_dev.hpp"
#include "tk_plan.hpp"

#include <cfloat>

namespace tk {

__global__ __launch_bounds__(kBlock) void k_oc_ratio(int n, const uint8_t* __restrict__ fl,
                                                      const double* __restrict__ dc,
                                                      const double* __restrict__ dv,
                                                      double* __restrict__ ratio,
                                                      double* partials,
                                                      Scalars* sc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double m = -INFINITY;
  if (e < n && fl[e] == 0) {
    const double d = fmax(dv[e], 1e-20);
    const double r = fmax(-dc[e] / d, 0.0);
    ratio[e] = r;
    m = r;
  }
  block_partial_max(m, partials, 0);
}

__global__ __launch_bounds__(kBlock) void k_oc_normalize(int n, const uint8_t* __restrict__ fl,
                                                          double* __restrict__ ratio,
                                                          const Scalars* __restrict__ sc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || fl[e] != 0) return;
  const double scale = sc->ratio_scale;
  ratio[e] = (scale > 0.0) ? ratio[e] / scale : 1.0;
}

__global__ void k_oc_init(Scalars* sc, double target, int max_steps, double tol, int probes) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  oc_init(sc, target, max_steps, tol, probes);
}

__global__ __launch_bounds__(kBlock) void k_oc_candidate(int n, const uint8_t* __restrict__ fl,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ ratio,
                                                          double move, int use_best,
                                                          double* __restrict__ cand,
                                                          Scalars* sc) {
  if (!use_best && sc->oc_done) return;
  const double lam = use_best ? sc->best_lam : sc->lam;   // sc->lam: set by init / finalize
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint8_t f = fl[e];
  cand[e] = oc_candidate(f, f == 0 ? x[e] : 0.0, f == 0 ? ratio[e] : 0.0, lam, move);
}

__global__ __launch_bounds__(kBlock) void k_oc_volume(Grid g, int nsten, const int* __restrict__ off,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ Hs,
                                                       const uint8_t* __restrict__ fl,
                                                       const double* __restrict__ cand,
                                                       double beta, double eta, double tb_eta,
                                                       double den, double* partials, Scalars* sc) {
  if (sc->oc_done) return;
  __shared__ int s_off[kMaxStencil];
  __shared__ double s_w[kMaxStencil];
  stage_stencil(nsten, off, w, s_off, s_w);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (e < g.nele && fl[e] == 0) {
    int iy, ix, iz;
    elem_coords(g, e, iy, ix, iz);
    const double f = stencil_gather(g, iy, ix, iz, nsten, s_off, s_w, cand) / Hs[e];
    double r, d;
    project(f, beta, eta, tb_eta, den, 0, r, d);
    v[0] = r;
  }
  block_partials<1>(v, partials);
}

// change = max |x_new - x| (free), vol_sum = sum rho (free), gray = #(0.05 < rho < 0.95) (free)
__global__ __launch_bounds__(kBlock) void k_design_stats(int n, const uint8_t* __restrict__ fl,
                                                          const double* __restrict__ xo,
                                                          const double* __restrict__ xn,
                                                          const double* __restrict__ rho,
                                                          double* partials,
                                                          Scalars* sc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  double m = -INFINITY;
  if (e < n && fl[e] == 0) {
    if (xo && xn) m = fabs(xn[e] - xo[e]);
    const double r = rho[e];
    v[0] = r;
    v[1] = (r > 0.05 && r < 0.95) ? 1.0 : 0.0;
  }
  block_sum<2>(v);
  m = block_max(m);
  if (threadIdx.x == 0) {
    const int G = gridDim.x;
    partials[blockIdx.x] = v[0];
    partials[G + blockIdx.x] = v[1];
    partials[2 * G + blockIdx.x] = m;
  }
}

// ---------------------------------------------------------------------------
void launch_oc_ratio(Plan& P, const double* dc, const double* dv, double* ratio, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_ratio, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele, P.eflags,
                     dc, dv, ratio, P.partials, P.sc);
  finalize(P, kFinRatioMax, nblocks(P.g.nele), s);
}

void launch_oc_normalize(Plan& P, double* ratio, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_normalize, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     P.eflags, ratio, P.sc);
}

void launch_oc_init(Plan& P, double target, int max_steps, double tol, int probes, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, s, P.sc, target, max_steps, tol, probes);
}

void launch_oc_candidate(Plan& P, const double* x, const double* ratio, double move, int use_best,
                         double* cand, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_candidate, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     P.eflags, x, ratio, move, use_best, cand, P.sc);
}

// One evaluation at sc->lam: the candidate (k_oc_candidate; forming it inside
// the tiled filter instead measured slower, the square root and division being
// repeated on the 2.5x larger halo-padded tile), its free-element projected
// volume, and the bisection update (kFinOcVolume).  All three return at once
// after the bisection has stopped.
void launch_oc_step(Plan& P, const double* x, const double* ratio, double move, double* cand,
                    double beta, double eta, double tb_eta, double den, hipStream_t s) {
  launch_oc_candidate(P, x, ratio, move, 0, cand, s);
  if (P.ftiled) {
    tiled_oc_volume(P, cand, beta, eta, tb_eta, den, s);
    return;
  }
  hipLaunchKernelGGL(k_oc_volume, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g, P.nsten,
                     P.sten_off, P.sten_w, P.Hs, P.eflags, cand, beta, eta, tb_eta, den,
                     P.partials, P.sc);
  finalize(P, kFinOcVolume, nblocks(P.g.nele), s);
}

void launch_design_stats(Plan& P, const double* x_old, const double* x_new, const double* rho,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_design_stats, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele,
                     P.eflags, x_old, x_new, rho, P.partials, P.sc);
  finalize(P, kFinStats, nblocks(P.g.nele), s);
}

}  // namespace tk
