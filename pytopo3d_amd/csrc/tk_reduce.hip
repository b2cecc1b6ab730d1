// Second phase of the grid reductions: one workgroup combines the per-block
// partials of the previous launch in a fixed order (deterministic) and applies
// the scalar update of the algorithm step that produced them (CG step lengths
// and convergence tests, the OC bisection bracket, compliance, statistics).
#include "tk_bisect.hpp"
#include "tk_plan.hpp"

#include <cfloat>

namespace tk {

__global__ __launch_bounds__(256) void k_finalize(int op, int G, const double* __restrict__ partials,
                                                  Scalars* sc, double a0, int i0) {
  if ((op == kFinAlpha || op == kFinUpdate || op == kFinRz || op == kFinUpdateFlex ||
       op == kFinRzPR) && sc->done)
    return;
  if (op == kFinOcVolume && sc->oc_done) return;
  double s[3] = {0.0, 0.0, 0.0};
  double m = -INFINITY;
  const bool two = (op == kFinNorms || op == kFinStats || op == kFinUpdateFlex);
  if (op == kFinRatioMax) {
    for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmax(m, partials[b]);
  } else {
#pragma unroll 4
    for (int b = threadIdx.x; b < G; b += blockDim.x) {
      s[0] += partials[b];
      if (two) s[1] += partials[G + b];
      if (op == kFinStats) m = fmax(m, partials[2 * G + b]);
    }
  }
  block_sum<3>(s);
  m = block_max(m);
  if (threadIdx.x != 0) return;
  switch (op) {
    case kFinSum0: sc->sum0 = s[0]; break;
    case kFinSum1: sc->sum1 = s[0]; break;
    case kFinNorm0: sc->norm0 = s[0]; break;
    case kFinNorm1: sc->norm1 = s[0]; break;
    case kFinC: sc->c = s[0]; break;
    case kFinAlpha:
      sc->pq = s[0];
      if (!(s[0] > 0.0)) {   // breakdown (or NaN): stop the iteration
        sc->alpha = 0.0;
        sc->done = 1;
        sc->stalled = 1;
      } else {
        sc->alpha = sc->rz / s[0];
      }
      break;
    case kFinNorms:   // s0 = b.b, s1 = r.r; a0 = tol^2, i0 = maxit
      sc->bb = s[0];
      sc->rr = s[1];
      sc->rr0 = s[1];
      sc->zs = sqrt(s[1]);
      sc->tol2 = a0;
      sc->maxit = i0;
      sc->iter = 0;
      sc->stalled = 0;
      sc->done = (s[1] <= a0 * s[0]) ? 1 : 0;
      if (!(s[1] == s[1])) { sc->done = 1; sc->stalled = 1; }
      break;
    case kFinUpdate:
      sc->zs = sqrt(sc->rr);   // the norm the fused update scaled the FP32 residual by
      sc->rr = s[0];
      sc->iter += 1;
      if (s[0] <= sc->tol2 * sc->bb) sc->done = 1;
      if (sc->iter >= sc->maxit) sc->done = 1;
      if (!(s[0] == s[0])) { sc->done = 1; sc->stalled = 1; }
      break;
    case kFinRz:      // i0 = first
      sc->beta = i0 ? 0.0 : s[0] / sc->rz;
      sc->rz = s[0];
      if (!(s[0] > 0.0)) { sc->done = 1; sc->stalled = (s[0] == 0.0) ? 0 : 1; }
      break;
    case kFinUpdateFlex:   // kFinUpdate, and s1 = q.z for the next beta
      sc->zs = sqrt(sc->rr);
      sc->rr = s[0];
      sc->qz = s[1];
      sc->iter += 1;
      if (s[0] <= sc->tol2 * sc->bb) sc->done = 1;
      if (sc->iter >= sc->maxit) sc->done = 1;
      if (!(s[0] == s[0])) { sc->done = 1; sc->stalled = 1; }
      break;
    case kFinRzPR: {   // s0 = r.z (i0 bit 1: of the scaled vectors, times zs^2); i0 bit 0: first
      const double rz = (i0 & 2) ? s[0] * sc->zs * sc->zs : s[0];
      // flexible (Polak-Ribiere) beta: (r.z - r.z_old) / rz_old, where
      // r.z_old = rz_old - alpha q_old.z_old
      sc->beta = (i0 & 1) ? 0.0 : (rz - (sc->rz - sc->alpha * sc->qz)) / sc->rz;
      sc->rz = rz;
      if (!(rz > 0.0)) { sc->done = 1; sc->stalled = (rz == 0.0) ? 0 : 1; }
      break;
    }
    case kFinRatioMax: sc->ratio_scale = m; break;
    case kFinOcVolume:   // the candidate was formed at sc->lam
      oc_record(sc, s[0]);
      break;
    case kFinStats:
      sc->vol_sum = s[0];
      sc->gray_cnt = s[1];
      sc->change = m;
      break;
    default: break;
  }
}

void finalize(Plan& P, int op, int G, hipStream_t s, double a0, int i0) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, op, G, P.partials, P.sc, a0, i0);
}

}  // namespace tk
