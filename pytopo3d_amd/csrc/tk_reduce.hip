// Second phase of the grid reductions: one workgroup combines the per-block
// partials of the previous launch in a fixed order (deterministic) and applies
// the scalar update of the algorithm step that produced them (CG step lengths
// and convergence tests, the OC bisection bracket, compliance, statistics).
#include "tk_bisect.hpp"
#include "tk_plan.hpp"

#include <cfloat>

namespace tk {

constexpr int kFinThreadsMax = 1024;
constexpr int kFinLoads = 8;   // independent loads in flight per thread

// blockDim.x: 256, or 1024 for more than 256 * kFinLoads partials (one batch of
// loads per thread either way up to 8192 partials; a 1024-thread block costs
// ~1 us more to run, which the solver's per-iteration reductions would pay)
__global__ __launch_bounds__(kFinThreadsMax) void k_finalize(int op, int G,
                                                          const double* __restrict__ partials,
                                                          Scalars* sc, double a0, int i0) {
  if ((op == kFinAlpha || op == kFinUpdate || op == kFinRz || op == kFinUpdateFlex ||
       op == kFinRzPR) && sc->done)
    return;
  if (op == kFinOcVolume && sc->oc_done) return;
  double s[3] = {0.0, 0.0, 0.0};
  double m = -INFINITY;
  const bool two = (op == kFinNorms || op == kFinStats || op == kFinUpdateFlex);
  // The partials were just written by other compute units, so every load is a
  // full memory round trip: each thread keeps kFinLoads independent loads in
  // flight (into separate accumulators, combined in a fixed order).
  {
    double a[kFinLoads], c[kFinLoads], x[kFinLoads];
#pragma unroll
    for (int j = 0; j < kFinLoads; ++j) {
      a[j] = c[j] = 0.0;
      x[j] = -INFINITY;
    }
    const bool is_max = op == kFinRatioMax, stats = op == kFinStats;
    const int T = blockDim.x;
    for (int b0 = threadIdx.x; b0 < G; b0 += kFinLoads * T) {
      // unconditional loads at clamped indices (G >= 1), masked afterwards
      double p[kFinLoads], q[kFinLoads], r[kFinLoads];
#pragma unroll
      for (int j = 0; j < kFinLoads; ++j) {
        const int b = min(b0 + j * T, G - 1);
        p[j] = partials[b];
        q[j] = two ? partials[G + b] : 0.0;
        r[j] = stats ? partials[2 * G + b] : -INFINITY;
      }
#pragma unroll
      for (int j = 0; j < kFinLoads; ++j) {
        const bool in = b0 + j * T < G;
        if (is_max) x[j] = fmax(x[j], in ? p[j] : -INFINITY);
        else a[j] += in ? p[j] : 0.0;
        c[j] += in ? q[j] : 0.0;
        x[j] = fmax(x[j], in ? r[j] : -INFINITY);
      }
    }
#pragma unroll
    for (int j = 0; j < kFinLoads; ++j) {
      s[0] += a[j];
      s[1] += c[j];
      m = fmax(m, x[j]);
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

void finalize_raw(int op, int G, const double* partials, Scalars* sc, hipStream_t s) {
  const int threads = G > 256 * kFinLoads ? kFinThreadsMax : 256;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(threads), 0, s, op, G, partials, sc, 0.0, 0);
}

void finalize(Plan& P, int op, int G, hipStream_t s, double a0, int i0, const double* partials) {
  const int threads = G > 256 * kFinLoads ? kFinThreadsMax : 256;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(threads), 0, s, op, G,
                     partials ? partials : P.partials, P.sc, a0, i0);
}

}  // namespace tk
