// Multiplier bisection of the projected-volume OC update
// (pytopo3d/utils/oc_update.py:76-124) with decisions taken from evaluations
// already made.
//
// The reference halves [0, 1e12] at lambda_mid = (lo + hi) / 2, forms the
// candidate, measures its projected free volume V and raises lo when
// V > target, else lowers hi, keeping the midpoint with the smallest
// |V - target| (the first one visited on ties) and stopping on the relative
// interval or the step limit.  V is non-increasing in lambda: the candidate
// x sqrt(max(ratio / (lambda + 1e-30), 1e-30)) clipped to the move limits and
// [0, 1], the filter with non-negative weights, the projection and the sum are
// all monotone.  So one evaluation at any lambda decides every midpoint on its
// side:
//   m < ka (V(ka) > target): V(m) >= V(ka) > target, lo = m;
//   m > kb (V(kb) <= target): V(m) <= V(kb) <= target, hi = m.
// oc_advance walks the reference's midpoints in its order with its arithmetic,
// so lo, hi, the step count and the stop are exactly the plain bisection's, and
// halts at the first midpoint inside (ka, kb), which must be measured.
// oc_next_probe chooses where the next evaluation goes: that midpoint, or a
// probe meant to pull ka and kb tight around the crossing (the previous call's
// multiplier, a Newton step in log lambda with the previous call's slope, or a
// search outwards when there is none, then two-sided secant steps in
// log lambda), after which most remaining midpoints are decided unmeasured.
//
// The kept midpoint.  The midpoints visited above the target increase (each
// becomes lo) and their errors V - target do not increase; those below
// decrease (each becomes hi) and their errors target - V do not increase.  So
// the reference keeps the final lo or the final hi, whichever has the smaller
// error (the one visited first on a tie), unless the previous midpoint on that
// side has exactly the same volume (a plateau: the first of equal ones is
// kept).  The bisection therefore tracks, per side, the last two midpoints
// with their volume when measured or its bound when implied; after the last
// step it measures what the choice still needs (at most three evaluations:
// the final lo and hi when a bound cannot decide between them, and the
// winner's predecessor when its bound does not separate it strictly).  A
// plateau at the end, an exact tie across the sides or a NaN volume makes the
// host repeat the bisection measuring every midpoint (probes off), which keeps
// the smallest error directly as the reference does.
#pragma once

#include <cfloat>
#include <cmath>

#include "tk_common.hpp"

namespace tk {

constexpr int kOcMaxProbes = 24;   // probe evaluations at most
constexpr int kOcMaxChecks = 3;    // evaluations after the last bisection step at most

// Side slots: 0 = final lo, 1 = the midpoint above the target visited before
// it, 2 = final hi, 3 = the midpoint below the target visited before it.
enum { kSlotNone = 0, kSlotBound = 1, kSlotExact = 2 };

__host__ __device__ inline void oc_next_probe(Scalars* sc);
__host__ __device__ inline void oc_finish(Scalars* sc);

// Start a bisection (probes: 0 measures every midpoint, the plain bisection).
// sc->lam_prev is kept from the previous bisection of the plan.
__host__ __device__ inline void oc_init(Scalars* sc, double target, int max_steps, double tol,
                                        int probes) {
  sc->lo = 0.0;
  sc->hi = 1.0e12;
  sc->best_err = INFINITY;
  sc->best_lam = 0.5e12;
  sc->target = target;
  sc->steps = 0;
  sc->oc_done = 0;
  sc->max_steps = max_steps;
  sc->tol = tol;
  sc->ka = 0.0;
  sc->kva = INFINITY;
  sc->kb = INFINITY;
  sc->kvb = -INFINITY;
  sc->gfac = 2.0;
  sc->evals = 0;
  sc->spec = probes;
  sc->nspec = 0;
  sc->redo = 0;
  sc->phase = 0;
  sc->check = -1;
  for (int k = 0; k < 4; ++k) {
    sc->slot_state[k] = kSlotNone;
    sc->slot_lam[k] = 0.0;
    sc->slot_v[k] = 0.0;
    sc->slot_step[k] = 0;
  }
  oc_next_probe(sc);
}

// Visit midpoint m on one side: shift the side's slots.
__host__ __device__ inline void oc_visit(Scalars* sc, bool above, double m, double v, int state) {
  const int s0 = above ? 0 : 2;
  sc->slot_state[s0 + 1] = sc->slot_state[s0];
  sc->slot_lam[s0 + 1] = sc->slot_lam[s0];
  sc->slot_v[s0 + 1] = sc->slot_v[s0];
  sc->slot_step[s0 + 1] = sc->slot_step[s0];
  sc->slot_state[s0] = state;
  sc->slot_lam[s0] = m;
  sc->slot_v[s0] = v;
  sc->slot_step[s0] = sc->steps;
}

// Take every bisection step whose decision is known (measured at that very
// multiplier, or implied by ka / kb).
__host__ __device__ inline void oc_advance(Scalars* sc) {
  while (sc->phase == 0) {
    const double m = 0.5 * (sc->lo + sc->hi);
    bool above;
    if (m == sc->ka || m == sc->kb) {
      const double Vm = (m == sc->ka) ? sc->kva : sc->kvb;
      // the reference's running choice (strict <): the answer when every
      // midpoint is measured
      const double err = fabs(Vm - sc->target);
      if (err < sc->best_err) {
        sc->best_err = err;
        sc->best_lam = m;
      }
      above = Vm > sc->target;
      oc_visit(sc, above, m, Vm, kSlotExact);
    } else if (m < sc->ka) {
      above = true;
      oc_visit(sc, true, m, sc->kva, kSlotBound);    // V(m) >= V(ka)
    } else if (m > sc->kb) {
      above = false;
      oc_visit(sc, false, m, sc->kvb, kSlotBound);   // V(m) <= V(kb)
    } else {
      break;   // inside (ka, kb): needs an evaluation
    }
    if (above) sc->lo = m;
    else sc->hi = m;
    sc->steps += 1;
    const double rel = (sc->hi - sc->lo) / fmax(sc->hi + sc->lo, DBL_MIN);
    if (rel < sc->tol || sc->steps >= sc->max_steps) sc->phase = 1;
  }
  if (sc->phase == 1) oc_finish(sc);
}

// After the last step: measure what the choice still needs, or choose.  The
// errors are exact where measured, lower bounds where implied (above:
// V >= the bound, below: V <= the bound).  A bound is measured only when it
// could change the decision: first between the final lo and hi, then the
// winner's predecessor, which must have a strictly larger error.
__host__ __device__ inline void oc_finish(Scalars* sc) {
  if (!sc->spec) {   // every midpoint was measured: the running choice stands
    sc->oc_done = 1;
    sc->lam_prev = sc->best_lam;
    return;
  }
  const int* st = sc->slot_state;
  const double* v = sc->slot_v;
  const double t = sc->target;
  auto measure = [&](int k) {
    sc->check = k;
    sc->lam = sc->slot_lam[k];
  };
  const bool ha = st[0] != kSlotNone, hb = st[2] != kSlotNone;
  const double ea = ha ? v[0] - t : INFINITY;   // exact, or a lower bound
  const double eb = hb ? t - v[2] : INFINITY;
  int win;   // 0: final lo, 2: final hi
  if (ha && hb) {
    const bool xa = st[0] == kSlotExact, xb = st[2] == kSlotExact;
    if (xa && xb) {
      if (ea == eb) {   // an exact tie across the sides: settle it by measuring all
        sc->redo = 1;
        sc->oc_done = 1;
        return;
      }
      win = ea < eb ? 0 : 2;
    } else if (xa && eb > ea) {
      win = 0;
    } else if (xb && ea > eb) {
      win = 2;
    } else {
      // measure the bound that may be the smaller error
      measure(!xa && (xb || ea <= eb) ? 0 : 2);
      return;
    }
  } else {
    win = ha ? 0 : 2;
    if (st[win] == kSlotBound) {
      measure(win);
      return;
    }
  }
  // the winner's predecessor on its side: strictly larger error
  const int p = win + 1;
  if (st[p] == kSlotBound) {
    const bool separated = (win == 0) ? (v[p] > v[0]) : (v[p] < v[2]);
    if (!separated) {
      measure(p);
      return;
    }
  } else if (st[p] == kSlotExact && v[p] == v[win]) {
    // a plateau: the kept midpoint may lie further back
    sc->redo = 1;
    sc->oc_done = 1;
    return;
  }
  sc->best_lam = sc->slot_lam[win];
  sc->best_err = (win == 0) ? ea : eb;
  sc->oc_done = 1;
  sc->lam_prev = sc->best_lam;
  if (sc->ka > 0.0 && sc->kb < INFINITY && sc->kb > sc->ka)
    sc->slope_prev = (sc->kva - sc->kvb) / log(sc->kb / sc->ka);
}

// Record an evaluation V at sc->lam and place the next one.
__host__ __device__ inline void oc_record(Scalars* sc, double V) {
  const double lp = sc->lam;
  sc->V = V;
  sc->evals += 1;
  if (sc->spec && !(V == V)) {   // NaN: no monotone reasoning; measure every midpoint
    sc->redo = 1;
    sc->oc_done = 1;
    return;
  }
  if (sc->phase == 1) {
    sc->slot_v[sc->check] = V;
    sc->slot_state[sc->check] = kSlotExact;
    oc_finish(sc);
    return;
  }
  if (V > sc->target) {
    if (lp > sc->ka) {
      sc->ka = lp;
      sc->kva = V;
    }
  } else if (lp < sc->kb) {
    sc->kb = lp;
    sc->kvb = V;
  }
  oc_advance(sc);
  if (sc->phase == 0) oc_next_probe(sc);
}

__host__ __device__ inline void oc_next_probe(Scalars* sc) {
  const double m = 0.5 * (sc->lo + sc->hi);   // the open midpoint, in (ka, kb)
  double p = m;
  if (sc->spec && sc->nspec < kOcMaxProbes) {
    const double a = fmax(sc->lo, sc->ka), b = fmin(sc->hi, sc->kb);
    const double w = (b - a) / (a + b);       // relative half-width of the open bracket
    // below ~2 tol only a couple of midpoints remain: measure them
    if (w > 2.0 * sc->tol) {
      const bool have_a = sc->ka > 0.0, have_b = sc->kb < INFINITY;
      double est = m, e = 0.0;
      if (have_a && have_b) {
        if (sc->kb > 2.0 * sc->ka) {
          est = sqrt(sc->ka * sc->kb);          // far apart: halve in log lambda
        } else {
          // secant in log lambda; its error shrinks like w^2, so the probe
          // goes that far past the estimate, on the side of the farther
          // bound, to land just inside it
          const double t = (sc->kva - sc->target) / (sc->kva - sc->kvb);
          est = sc->ka * pow(sc->kb / sc->ka, t);
          e = fmax(0.25 * w * w, 0.25 * sc->tol);
          const bool low_far = est * est > sc->ka * sc->kb;
          est *= low_far ? (1.0 - e) : (1.0 + e);
        }
      } else if (sc->nspec == 0 && sc->lam_prev > a && sc->lam_prev < b) {
        est = sc->lam_prev;                     // the previous design's multiplier
        sc->gfac = 1.05;
      } else if (sc->nspec == 1 && sc->slope_prev > 0.0 && (have_a != have_b)) {
        // one side known, from the previous multiplier: a Newton step in
        // log lambda with the previous slope, overshot by a quarter so that it
        // lands past the crossing
        const double l0 = have_a ? sc->ka : sc->kb, v0 = have_a ? sc->kva : sc->kvb;
        const double step = (v0 - sc->target) / sc->slope_prev;   // > 0 above, < 0 below
        est = l0 * exp(1.25 * step);
      } else if (have_a) {
        est = sc->ka * sc->gfac;                // search upwards, doubling in log lambda
        sc->gfac *= sc->gfac;
      } else if (have_b) {
        est = sc->kb / sc->gfac;                // search downwards
        sc->gfac *= sc->gfac;
      }
      if (est > a && est < b) p = est;
      // a probe about at the open midpoint: measure the midpoint itself
      if (fabs(p - m) <= fmax(e, sc->tol) * m) p = m;
    }
    if (p != m) sc->nspec += 1;
  }
  sc->lam = p;
}

}  // namespace tk
