// Host check of the OC multiplier bisection (pytopo3d_amd/csrc/tk_bisect.hpp,
// the same functions the device finalize runs): with probes on, the kept
// multiplier, its error and the step count must equal those of the plain
// bisection (every midpoint measured, pytopo3d/utils/oc_update.py:76-124) on
// seeded monotone volume functions, and each call must stay within
// max_steps + kOcMaxProbes + kOcMaxChecks evaluations.  Prints the mean
// evaluation counts.
//
// The volume function mimics the OC one: per element the clipped candidate
// x sqrt(max(r / (lam + 1e-30), 1e-30)) within the move limits and [0, 1],
// passed through a monotone smooth map and summed in a fixed order, so it is
// non-increasing in lam.  Cases cover smooth fields, many identical elements
// (plateaus and kinks), a target above and below the reachable range, one
// element, and a sequence of slowly changing designs (the previous multiplier
// as the first probe).
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "tk_bisect.hpp"

namespace {

struct Field {
  std::vector<double> x, r;
  double move = 0.2;
  int shape = 0;
};

double volume(const Field& F, double lam) {
  double v = 0.0;
  for (size_t i = 0; i < F.x.size(); ++i) {
    const double xe = F.x[i];
    double c = xe * std::sqrt(std::fmax(F.r[i] / (lam + 1e-30), 1e-30));
    c = std::fmax(xe - F.move, std::fmin(xe + F.move, c));
    c = std::fmin(std::fmax(c, 0.0), 1.0);
    if (F.shape == 1) c = c * c;
    if (F.shape == 2) c = c / (1.0 + c);
    v += c;
  }
  return v;
}

struct Result {
  double lam, err;
  int steps, evals, redo;
};

Result run(const Field& F, double target, int probes, tk::Scalars& sc, int max_steps, double tol,
           bool& cap_ok) {
  tk::oc_init(&sc, target, max_steps, tol, probes);
  int n = 0;
  const int cap = max_steps + tk::kOcMaxProbes + tk::kOcMaxChecks;
  while (!sc.oc_done && n < cap + 5) {
    tk::oc_record(&sc, volume(F, sc.lam));
    ++n;
  }
  cap_ok = sc.oc_done && n <= cap;
  return {sc.best_lam, sc.best_err, sc.steps, sc.evals, sc.redo};
}

}  // namespace

int main() {
  std::mt19937_64 rng(20240613);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  int cases = 0, bad = 0, redos = 0;
  long long ev_plain = 0, ev_probe = 0;
  int kcases[8] = {0}, kredo[8] = {0};
  long long kplain[8] = {0}, kprobe[8] = {0};
  tk::Scalars sc{};       // lam_prev carries over between calls, as on a plan
  tk::Scalars plain{};
  for (int c = 0; c < 400; ++c) {
    Field F;
    const int kind = c % 8;
    const int n = (kind == 6) ? 1 : 50 + (int)(U(rng) * 400);
    F.shape = c % 3;
    F.move = (c % 5 == 0) ? 0.05 : 0.2;
    const double p = 0.5 + 3.0 * U(rng);
    F.x.resize(n);
    F.r.resize(n);
    for (int i = 0; i < n; ++i) {
      F.x[i] = (kind == 4) ? 0.3 : U(rng);
      F.r[i] = (kind == 4) ? 1.0 : std::pow(U(rng), p);
    }
    double rmax = 0.0;
    for (double v : F.r) rmax = std::fmax(rmax, v);
    for (double& v : F.r) v = rmax > 0.0 ? v / rmax : 1.0;
    const double vmin = volume(F, 1e12), vmax = volume(F, 1e-9);
    // a sequence of slowly changing designs: the multiplier moves a little
    const int seq = (kind == 7) ? 6 : 1;
    for (int s = 0; s < seq; ++s) {
      double target;
      if (kind == 5) target = (c & 1) ? vmax + 1.0 : vmin - 1.0;
      else target = vmin + (0.1 + 0.8 * U(rng)) * (vmax - vmin);
      if (s > 0)
        for (double& v : F.r) v = std::fmin(1.0, v * (1.0 + 0.02 * (U(rng) - 0.5)));
      for (int tolk = 0; tolk < 2; ++tolk) {
        const double tol = tolk ? 1e-8 : 1e-10;
        const int max_steps = tolk ? 80 : 60;
        bool cap1 = false, cap2 = false;
        const Result a = run(F, target, 0, plain, max_steps, tol, cap1);
        Result b = run(F, target, 1, sc, max_steps, tol, cap2);
        const int probe_evals = b.evals;
        ++kcases[kind];
        kplain[kind] += a.evals;
        if (b.redo) {
          ++redos;
          ++kredo[kind];
          bool cap3 = false;
          b = run(F, target, 0, sc, max_steps, tol, cap3);
          b.evals += probe_evals;
          cap2 = cap2 && cap3;
        }
        ++cases;
        ev_plain += a.evals;
        ev_probe += b.evals;
        kprobe[kind] += b.evals;
        if (a.lam != b.lam || a.err != b.err || a.steps != b.steps || !cap1 || !cap2) {
          ++bad;
          if (bad <= 10)
            std::printf("MISMATCH case %d seq %d tol %g: plain lam %.17g err %.3g steps %d | "
                        "probes lam %.17g err %.3g steps %d evals %d cap %d%d\n",
                        c, s, tol, a.lam, a.err, a.steps, b.lam, b.err, b.steps, b.evals, cap1,
                        cap2);
        }
      }
    }
  }
  const char* names[8] = {"smooth", "smooth", "smooth", "smooth", "identical elements",
                          "target out of range", "one element", "design sequence"};
  for (int k = 0; k < 8; ++k)
    std::printf("  %-20s cases %4d repeated %3d evaluations plain %.2f probes %.2f\n", names[k],
                kcases[k], kredo[k], (double)kplain[k] / kcases[k], (double)kprobe[k] / kcases[k]);
  std::printf("cases %d mismatches %d repeated %d mean evaluations plain %.2f probes %.2f\n", cases,
              bad, redos, (double)ev_plain / cases, (double)ev_probe / cases);
  return bad ? 1 : 0;
}
