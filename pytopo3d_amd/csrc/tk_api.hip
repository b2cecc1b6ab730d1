// C ABI of libpytopo3d_hip.so (declared in include/pytopo3d_hip.h): plan
// lifecycle, argument validation with the reference's error messages, and the
// entry points that enqueue the kernels of the other translation units.
#include "pytopo3d_hip.h"
#include "tk_bisect.hpp"
#include "tk_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace tk {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

int apply_fine(Plan& P, int prec, const void* u, void* y, hipStream_t s);
int smooth_step_fine(Plan& P, int prec, const void* b, const void* x, void* xo, void* d,
                     double c1, double c2, hipStream_t s);
void launch_oc_init(Plan& P, double target, int max_steps, double tol, int probes, hipStream_t s);
void launch_chain_weights(Plan& P, const double* gc, const double* gv, const double* drho,
                          double* wc, double* wv, hipStream_t s);
void launch_boundary_conditions(Plan& P, const double* ff, const uint8_t* sm, const uint8_t* ob,
                                const uint8_t* pz, double* F, uint8_t* fixed, uint8_t* flags,
                                hipStream_t s);

// Largest eigenvalue of diag(KE)^-1/2 KE diag(KE)^-1/2 (cyclic Jacobi rotations).
static double lambda_bound(const double* KE) { return scaled_lambda_max(KE); }

template <typename T>
static int dmalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  TK_HIP(hipMalloc((void**)p, sizeof(T) * n));
  return 0;
}

static void free_plan(Plan* P) {
  if (P->cstream) (void)hipStreamSynchronize(P->cstream);   // a factorisation may be running
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(P->KE64); f(P->KE32); f(P->sten_off); f(P->sten_w); f(P->Hs); f(P->eflags); f(P->fcols);
  f(P->sec64_dev); f(P->sec32_dev); f(P->kemf);
  for (int l = 0; l < kMaxLevels; ++l) {
    Level& L = P->lv[l];
    f(L.fixed);
    f(L.f64.s); f(L.f64.dinv); f(L.f64.x); f(L.f64.xa); f(L.f64.b); f(L.f64.r); f(L.f64.d);
    f(L.f32.s); f(L.f32.dinv); f(L.f32.x); f(L.f32.xa); f(L.f32.b); f(L.f32.r); f(L.f32.d);
  }
  f(P->coarse_inv64); f(P->coarse_inv32); f(P->coarse_L);
  f(P->r64); f(P->p64); f(P->q64); f(P->x64);
  f(P->r32); f(P->p32); f(P->q32); f(P->x32); f(P->pring); f(P->alpha_hist);
  f(P->ework0); f(P->ework1); f(P->ework2); f(P->ework3);
  f(P->partials); f(P->sc); f(P->pc_sync);
  f(P->gal_tab64); f(P->gal_tab32); f(P->gal_dw); f(P->gal_sh64); f(P->gal_sh32);
  if (P->sc_host) (void)hipHostFree(P->sc_host);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < Plan::kChunkSizes; ++c)
        for (int d = 0; d < 2; ++d)
          if (P->gexec[a][b][c][d]) (void)hipGraphExecDestroy(P->gexec[a][b][c][d]);
  if (P->gstream) (void)hipStreamDestroy(P->gstream);
  if (P->cstream) (void)hipStreamDestroy(P->cstream);
  if (P->cev_in) (void)hipEventDestroy(P->cev_in);
  if (P->cev_out) (void)hipEventDestroy(P->cev_out);
  delete P;
}

template <typename T>
static int alloc_level_buf(LevelBuf<T>& B, const Grid& g) {
  if (dmalloc(&B.s, g.nele) || dmalloc(&B.dinv, g.ndof) || dmalloc(&B.x, g.ndof) ||
      dmalloc(&B.xa, g.ndof) || dmalloc(&B.b, g.ndof) || dmalloc(&B.r, g.ndof) ||
      dmalloc(&B.d, g.ndof))
    return 2;
  return 0;
}

static int create(Plan*& out, int nelx, int nely, int nelz, const double* KE, double rmin,
                  const tk_options* o) {
  if (nelx < 1 || nely < 1 || nelz < 1) {
    set_error("nelx, nely, and nelz must all be positive");
    return 1;
  }
  if (!(std::isfinite(rmin) && rmin > 0.0)) {
    set_error("rmin must be finite and positive, got " + std::to_string(rmin));
    return 1;
  }
  const long long nn = (long long)(nelx + 1) * (nely + 1) * (nelz + 1);
  if (3 * nn >= (1LL << 31) - 1024) {
    set_error("grid too large: the DOF count must fit in 31 bits");
    return 1;
  }
  if (!KE) {
    set_error("KE must be a host pointer to 576 doubles");
    return 1;
  }
  for (int i = 0; i < 24; ++i) {
    if (!(KE[i * 25] > 0.0) || !std::isfinite(KE[i * 25])) {
      set_error("KE must have a positive finite diagonal");
      return 1;
    }
    for (int j = 0; j < 24; ++j)
      if (!std::isfinite(KE[i * 24 + j]) ||
          std::fabs(KE[i * 24 + j] - KE[j * 24 + i]) > 1e-10 * std::fabs(KE[i * 25])) {
        set_error("KE must be finite and symmetric");
        return 1;
      }
  }
  Plan* P = new Plan();
  if (o) {
    if (o->smooth_degree > 0) P->opt.smooth_degree = o->smooth_degree;
    if (o->coarse_degree > 0) P->opt.coarse_degree = o->coarse_degree;
    if (o->coarse_cheb_lo > 0.0) P->opt.coarse_cheb_lo = o->coarse_cheb_lo;
    if (o->mid_degree > 0) P->opt.mid_degree = o->mid_degree;
    if (o->mid_cheb_lo > 0.0) P->opt.mid_cheb_lo = o->mid_cheb_lo;
    if (o->cheb_lo > 0.0) P->opt.cheb_lo = o->cheb_lo;
    if (o->cheb_hi > 0.0) P->opt.cheb_hi = o->cheb_hi;
    if (o->max_levels > 0) P->opt.max_levels = o->max_levels < kMaxLevels ? o->max_levels : kMaxLevels;
    if (o->coarse_max_dof > 0)
      P->opt.coarse_max_dof = o->coarse_max_dof < kCoarseMaxDof ? o->coarse_max_dof : kCoarseMaxDof;
    if (o->no_graphs == 0 || o->no_graphs == 1) P->opt.no_graphs = o->no_graphs;
    if (o->filter_kind == 0 || o->filter_kind == 1) P->opt.filter_kind = o->filter_kind;
    if (o->check_every > 0) P->opt.check_every = o->check_every;
    if (o->operator_kind >= 0 && o->operator_kind <= 2) P->opt.operator_kind = o->operator_kind;
    if (o->direction_slots != 0 && (o->direction_slots < 2 || o->direction_slots > kRingMax)) {
      delete P;
      set_error("direction_slots must be 0 or in [2, " + std::to_string(kRingMax) + "]");
      return 1;
    }
    P->opt.direction_slots = o->direction_slots;
    if (o->galerkin >= 0 && o->galerkin <= 2) P->opt.galerkin = o->galerkin;
    if (o->persistent_coarse == 1 || o->persistent_coarse == -1)
      P->opt.persistent_coarse = o->persistent_coarse == 1;
  }
  if (!(P->opt.cheb_lo < P->opt.cheb_hi) || !(P->opt.coarse_cheb_lo < P->opt.cheb_hi) ||
      !(P->opt.mid_cheb_lo < P->opt.cheb_hi)) {
    delete P;
    set_error("cheb_lo, coarse_cheb_lo and mid_cheb_lo must be below cheb_hi");
    return 1;
  }
  {
    // the coarsest grid is factorised in LDS: n(n+1)/2 doubles must fit one workgroup
    int dev = 0, shm = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&shm, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        shm <= 0)
      shm = 65536;
    int nmax = 1;
    while ((long long)(nmax + 1) * (nmax + 2) / 2 * 8 + 256 <= shm) ++nmax;
    if (P->opt.coarse_max_dof > nmax) P->opt.coarse_max_dof = nmax;
    P->lds_limit = shm;
    if (hipDeviceGetAttribute(&P->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        P->ncu <= 0)
      P->ncu = 256;
  }
  P->g = make_grid(nelx, nely, nelz);
  std::memcpy(P->KE, KE, sizeof(double) * 576);
  P->lam_bound = lambda_bound(KE);
  operator_setup(*P);
  // the element kernel's smoothing forms D^-1 from KE's corner-independent diagonal
  P->use_el = P->sector_ok && P->kd_ok && P->opt.operator_kind != 1;
  P->mfma = P->use_el && P->opt.operator_kind == 2 && P->kemf;
  P->rmin = rmin;
  // stencil: integer offsets with rmin - |d| > 1e-9, ordered (dz, dx, dy) ascending
  const int R = (int)std::ceil(rmin);
  std::vector<int> off;
  std::vector<double> w;
  if (R > 15) {
    delete P;
    set_error("rmin too large for the stencil filter (rmin <= 15)");
    return 1;
  }
  for (int dz = -R; dz <= R; ++dz)
    for (int dx = -R; dx <= R; ++dx)
      for (int dy = -R; dy <= R; ++dy) {
        const double wt = rmin - std::sqrt((double)(dx * dx + dy * dy + dz * dz));
        if (wt > 1e-9) {
          off.push_back((dy + 16) | ((dx + 16) << 8) | ((dz + 16) << 16));
          w.push_back(wt);
          P->reach = std::max(P->reach, std::max(std::abs(dy), std::max(std::abs(dx), std::abs(dz))));
        }
      }
  if ((int)off.size() > kMaxStencil) {
    delete P;
    set_error("rmin too large: filter stencil exceeds " + std::to_string(kMaxStencil) + " entries");
    return 1;
  }
  P->nsten = (int)off.size();
  P->R = R;
  // levels
  int nl = 1;
  P->lv[0].g = P->g;
  while (nl < P->opt.max_levels && P->lv[nl - 1].g.ndof > P->opt.coarse_max_dof) {
    const Grid& f = P->lv[nl - 1].g;
    if (f.nelx == 1 && f.nely == 1 && f.nelz == 1) break;
    P->lv[nl].g = make_grid((f.nelx + 1) / 2, (f.nely + 1) / 2, (f.nelz + 1) / 2);
    ++nl;
  }
  P->nlev = nl;
  P->ncoarse = P->lv[nl - 1].g.ndof;
  if (P->ncoarse > P->opt.coarse_max_dof && P->ncoarse > 24) {
    free_plan(P);
    set_error("coarsest multigrid grid too large (" + std::to_string(P->ncoarse) +
              " DOFs); raise max_levels");
    return 1;
  }
  // allocations
  const Grid& g = P->g;
  int rc = 0;
  rc |= dmalloc(&P->KE64, 576) | dmalloc(&P->KE32, 576) | dmalloc(&P->sten_off, P->nsten) |
        dmalloc(&P->sten_w, P->nsten) | dmalloc(&P->Hs, g.nele) | dmalloc(&P->eflags, g.nele);
  for (int l = 0; l < nl && !rc; ++l) {
    Level& L = P->lv[l];
    rc |= dmalloc(&L.fixed, L.g.nnode);
    rc |= alloc_level_buf(L.f64, L.g);
    rc |= alloc_level_buf(L.f32, L.g);
  }
  const int nc = P->ncoarse;
  rc |= dmalloc(&P->coarse_inv64, (size_t)nc * nc) | dmalloc(&P->coarse_inv32, (size_t)nc * nc) |
        dmalloc(&P->coarse_L, (size_t)nc * (nc + 1) / 2);
  rc |= dmalloc(&P->r64, g.ndof) | dmalloc(&P->p64, g.ndof) | dmalloc(&P->q64, g.ndof) |
        dmalloc(&P->x64, g.ndof);
  rc |= dmalloc(&P->r32, g.ndof) | dmalloc(&P->p32, g.ndof) | dmalloc(&P->q32, g.ndof) |
        dmalloc(&P->x32, g.ndof);
  {
    // slot count of the mixed solver's direction ring (allocated by its first
    // mixed solve)
    const long long slots = kRingBytes / (4LL * g.ndof);
    P->ring_m = P->opt.direction_slots > 0
                    ? P->opt.direction_slots
                    : (int)(slots < 2 ? 2 : (slots > kRingMax ? kRingMax : slots));
  }
  rc |= dmalloc(&P->ework0, g.nele) | dmalloc(&P->ework1, g.nele) | dmalloc(&P->ework2, g.nele) |
        dmalloc(&P->ework3, g.nele);
  const long long big = g.ndof > g.nele ? g.ndof : g.nele;
  P->partials_len = 4 * (nblocks(big) + 16);
  rc |= dmalloc(&P->partials, P->partials_len) | dmalloc(&P->sc, 1) | dmalloc(&P->pc_sync, kPcSyncWords);
  if (rc) {
    free_plan(P);
    return 2;
  }
  if (hipHostMalloc((void**)&P->sc_host, sizeof(Scalars)) != hipSuccess ||
      hipHostGetDevicePointer((void**)&P->sc_host_dev, P->sc_host, 0) != hipSuccess) {
    free_plan(P);
    set_error("hipHostMalloc failed");
    return 2;
  }
  if (hipStreamCreateWithFlags(&P->cstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&P->cev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->cev_out, hipEventDisableTiming) != hipSuccess) {
    free_plan(P);
    set_error("could not create the set-up stream");
    return 2;
  }
  std::vector<float> ke32(576);
  for (int i = 0; i < 576; ++i) ke32[i] = (float)KE[i];
  hipStream_t s = nullptr;
  Scalars zero{};
  std::vector<uint8_t> zeros((size_t)g.nnode, 0);
  hipError_t e = hipSuccess;
  e = e ? e : hipMemcpy(P->KE64, KE, sizeof(double) * 576, hipMemcpyHostToDevice);
  e = e ? e : hipMemcpy(P->KE32, ke32.data(), sizeof(float) * 576, hipMemcpyHostToDevice);
  e = e ? e : hipMemcpy(P->sten_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice);
  e = e ? e : hipMemcpy(P->sten_w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice);
  e = e ? e : hipMemcpy(P->sc, &zero, sizeof(Scalars), hipMemcpyHostToDevice);
  e = e ? e : hipMemset(P->eflags, 0, g.nele);
  e = e ? e : hipMemset(P->pc_sync, 0, sizeof(unsigned) * kPcSyncWords);
  for (int l = 0; l < nl && e == hipSuccess; ++l)
    e = hipMemset(P->lv[l].fixed, 0, P->lv[l].g.nnode);
  const size_t lds = sizeof(double) * (size_t)nc * (nc + 1) / 2;
  e = e ? e : coarse_setup_attr(lds);
  if (e != hipSuccess) {
    set_error(std::string("plan initialisation failed: ") + hipGetErrorString(e));
    free_plan(P);
    return 2;
  }
  if (P->opt.filter_kind == 1 || !filter_tiled_setup(*P, off, w)) P->ftiled = false;
  tail_setup(*P);
  // Galerkin operator on level 2 when asked for, and when that level is a
  // level of its own (not in the single-workgroup tail, not the coarsest) and
  // KE has the expected pattern
  if (P->opt.galerkin == 1 && P->use_el && P->nlev >= 4 &&
      (P->tail_level < 0 || P->tail_level > 2) && galerkin_setup(*P)) {
    P->gal_level = 2;
    const size_t ne = 8 * (size_t)P->lv[2].g.nele;
    if (dmalloc(&P->gal_sh64, ne) || dmalloc(&P->gal_sh32, ne)) {
      free_plan(P);
      return 2;
    }
  }
  pcycle_setup(*P);
  launch_filter_hs(*P, s);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string("filter setup failed: ") + hipGetErrorString(e));
    free_plan(P);
    return 2;
  }
  P->nfree = g.nele;
  out = P;
  return 0;
}

static int check_beta_eta(double beta, double eta) {
  if (!std::isfinite(beta) || beta <= 0) {
    set_error("beta must be positive, got " + std::to_string(beta));
    return 1;
  }
  if (!(0.0 < eta && eta < 1.0)) {
    set_error("eta must be between 0 and 1, got " + std::to_string(eta));
    return 1;
  }
  return 0;
}

static void proj_consts(double beta, double eta, double& tb, double& den) {
  tb = std::tanh(beta * eta);
  den = tb + std::tanh(beta * (1.0 - eta));
}

static int read_sc(Plan& P, hipStream_t s) {
  TK_HIP(hipMemcpyAsync(P.sc_host, P.sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  TK_HIP(hipStreamSynchronize(s));
  return 0;
}

// Enqueue OC evaluations in chunks until the device bisection has stopped.
// Every evaluation measures a midpoint (at most max_steps), is one of at most
// kOcMaxProbes probes or one of at most kOcMaxChecks checks after the last
// step, so the cap is never the stop.
static int oc_bisect(Plan& P, const double* x, double move, double beta, double eta, double tb,
                     double den, int max_steps, hipStream_t s) {
  const int cap = max_steps + kOcMaxProbes + kOcMaxChecks;
  int launched = 0, chunk = 12;
  for (;;) {
    const int n = std::min(chunk, cap - launched);
    for (int k = 0; k < n; ++k)
      launch_oc_step(P, x, P.ework0, move, P.ework1, beta, eta, tb, den, s);
    launched += n;
    if (read_sc(P, s)) return 2;
    if (P.sc_host->oc_done || launched >= cap) break;
    chunk = 8;
  }
  P.cnt.oc_steps += P.sc_host->evals;
  return 0;
}

}  // namespace tk

struct tk_plan {
  tk::Plan* p;
};

using tk::set_error;

#define TK_PLAN_OR_FAIL(plan)               \
  if (!(plan) || !(plan)->p) {              \
    set_error("null plan");                 \
    return 1;                               \
  }                                         \
  tk::Plan& P = *(plan)->p;

#define TK_LAUNCH_CHECK()                                                                \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess) {                                                              \
      set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_));         \
      return 2;                                                                          \
    }                                                                                    \
  } while (0)

extern "C" {

void tk_options_default(tk_options* o) {
  if (!o) return;
  tk::Options d;
  o->smooth_degree = d.smooth_degree;
  o->cheb_lo = d.cheb_lo;
  o->cheb_hi = d.cheb_hi;
  o->max_levels = d.max_levels;
  o->coarse_max_dof = d.coarse_max_dof;
  o->check_every = d.check_every;
  o->operator_kind = d.operator_kind;
  o->no_graphs = d.no_graphs;
  o->filter_kind = d.filter_kind;
  o->coarse_degree = d.coarse_degree;
  o->coarse_cheb_lo = d.coarse_cheb_lo;
  o->direction_slots = d.direction_slots;
  o->galerkin = d.galerkin;
  o->persistent_coarse = d.persistent_coarse;
  o->mid_degree = d.mid_degree;
  o->mid_cheb_lo = d.mid_cheb_lo;
}

int tk_plan_create(tk_plan** out, int nelx, int nely, int nelz, const double* KE, double rmin,
                   const tk_options* opt) {
  if (!out) {
    set_error("out must not be NULL");
    return 1;
  }
  *out = nullptr;
  tk::Plan* P = nullptr;
  const int rc = tk::create(P, nelx, nely, nelz, KE, rmin, opt);
  if (rc) return rc;
  *out = new tk_plan{P};
  return 0;
}

int tk_plan_destroy(tk_plan* plan) {
  if (!plan) return 0;
  if (plan->p) {
    (void)hipDeviceSynchronize();
    tk::free_plan(plan->p);
  }
  delete plan;
  return 0;
}

int tk_plan_set_masks(tk_plan* plan, const uint8_t* elem_flags, const uint8_t* node_fixed,
                      long long* nfree, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  hipStream_t s = (hipStream_t)stream;
  if (!elem_flags || !node_fixed) {
    set_error("elem_flags and node_fixed are required");
    return 1;
  }
  P.mask_gen++;   // invalidates the set-up a TK_SOLVE_REUSE would use
  TK_HIP(hipMemcpyAsync(P.eflags, elem_flags, P.g.nele, hipMemcpyDeviceToDevice, s));
  TK_HIP(hipMemcpyAsync(P.lv[0].fixed, node_fixed, P.g.nnode, hipMemcpyDeviceToDevice, s));
  for (int l = 0; l + 1 < P.nlev; ++l) tk::launch_coarsen_fixed(P.lv[l], P.lv[l + 1], s);
  tk::launch_count_free(P, s);
  TK_LAUNCH_CHECK();
  if (tk::read_sc(P, s)) return 2;
  P.nfree = (long long)P.sc_host->sum0;
  P.masks_set = true;
  if (nfree) *nfree = P.nfree;
  return 0;
}

int tk_build_boundary_conditions(tk_plan* plan, const double* force_field,
                                 const uint8_t* support_mask, const uint8_t* obstacle_mask,
                                 const uint8_t* protected_zone_mask, double* F,
                                 uint8_t* node_fixed, uint8_t* elem_flags, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!F && !node_fixed && !elem_flags) {
    set_error("at least one of F, node_fixed and elem_flags is required");
    return 1;
  }
  tk::launch_boundary_conditions(P, force_field, support_mask, obstacle_mask, protected_zone_mask,
                                 F, node_fixed, elem_flags, (hipStream_t)stream);
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_plan_levels(const tk_plan* plan, int* nlev, int* dims, int cap) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  const tk::Plan& P = *plan->p;
  if (nlev) *nlev = P.nlev;
  for (int l = 0; l < P.nlev && l < cap && dims; ++l) {
    dims[3 * l] = P.lv[l].g.nelx;
    dims[3 * l + 1] = P.lv[l].g.nely;
    dims[3 * l + 2] = P.lv[l].g.nelz;
  }
  return 0;
}

int tk_plan_stencil(const tk_plan* plan, int* offsets, double* weights, int cap, int* n) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  const tk::Plan& P = *plan->p;
  if (n) *n = P.nsten;
  if (!offsets && !weights) return 0;
  std::vector<int> off(P.nsten);
  std::vector<double> w(P.nsten);
  if (hipMemcpy(off.data(), P.sten_off, sizeof(int) * P.nsten, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(w.data(), P.sten_w, sizeof(double) * P.nsten, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("stencil copy failed");
    return 2;
  }
  for (int k = 0; k < P.nsten && k < cap; ++k) {
    if (offsets) {
      offsets[3 * k] = (off[k] & 255) - 16;
      offsets[3 * k + 1] = ((off[k] >> 8) & 255) - 16;
      offsets[3 * k + 2] = ((off[k] >> 16) & 255) - 16;
    }
    if (weights) weights[k] = w[k];
  }
  return 0;
}

int tk_plan_filter_info(const tk_plan* plan, int* tiled, int* halo, int* nsten) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  const tk::Plan& P = *plan->p;
  if (tiled) *tiled = P.ftiled ? 1 : 0;
  if (halo) *halo = P.ftiled ? P.fhalo : P.reach;
  if (nsten) *nsten = P.nsten;
  return 0;
}

int tk_filter_row_sums(tk_plan* plan, double* Hs, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  TK_HIP(hipMemcpyAsync(Hs, P.Hs, sizeof(double) * P.g.nele, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return 0;
}

int tk_filter_project(tk_plan* plan, const double* x, double beta, double eta, double* rho_f,
                      double* rho, double* drho, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (tk::check_beta_eta(beta, eta)) return 1;
  if (!x) {
    set_error("x is required");
    return 1;
  }
  double tb, den;
  tk::proj_consts(beta, eta, tb, den);
  tk::launch_filter_project(P, x, beta, eta, tb, den, rho_f, rho, drho, (hipStream_t)stream);
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_filter_transpose2(tk_plan* plan, const double* g_c, const double* g_v,
                         const double* drho, double* dc_dx, double* dv_dx, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!g_c || !drho || !dc_dx) {
    set_error("g_c, drho and dc_dx are required");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  tk::launch_chain_weights(P, g_c, g_v, drho, P.ework2, dv_dx ? P.ework3 : nullptr, s);
  tk::launch_filter_transpose2(P, P.ework2, dv_dx ? P.ework3 : nullptr, dc_dx, dv_dx, s);
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_apply_K(tk_plan* plan, int prec, const double* rho, double penal, const void* u, void* y,
               void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!u || !y || (rho && !(std::isfinite(penal) && penal > 0))) {
    set_error("u and y are required and penal must be positive");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  if (rho) {
    if (prec == TK_FP64) tk::launch_stiffness_scale<double>(P, rho, penal, P.lv[0].f64.s, s);
    else tk::launch_stiffness_scale<float>(P, rho, penal, P.lv[0].f32.s, s);
    P.last_penal = -1.0;
  } else if (P.last_penal < 0.0 || (prec == TK_MIXED_FP32 && P.last_prec == TK_FP64)) {
    set_error("rho == NULL reuses the scales of the last tk_solve in this precision, but there is none");
    return 1;
  }
  return tk::apply_fine(P, prec, u, y, s);
}

int tk_level_apply(tk_plan* plan, int prec, int level, const void* u, void* y, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!u || !y) {
    set_error("u and y are required");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  if (level < 0 || level >= P.nlev) {
    set_error("level must be in [0, " + std::to_string(P.nlev) + ")");
    return 1;
  }
  if (P.last_penal < 0.0 || prec != P.last_prec) {
    set_error("tk_level_apply needs a previous tk_solve or tk_prepare in this precision");
    return 1;
  }
  return tk::apply_level(P, prec, level, u, y, (hipStream_t)stream);
}

int tk_transfer_apply(tk_plan* plan, int prec, int level, int direction, const void* in,
                      void* out, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!in || !out) {
    set_error("in and out are required");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  if (level < 0 || level + 1 >= P.nlev) {
    set_error("level must be in [0, " + std::to_string(P.nlev - 1) + ")");
    return 1;
  }
  if (direction != 0 && direction != 1) {
    set_error("direction must be 0 (restriction) or 1 (prolongation)");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  const tk::Level& F = P.lv[level];
  const tk::Level& C = P.lv[level + 1];
  if (prec == TK_FP64) {
    if (direction == 0)
      tk::launch_restrict<double>(F, C, (const double*)in, (double*)out, s, nullptr);
    else
      tk::launch_prolong_add<double>(C, F, (const double*)in, (double*)out, s, nullptr);
  } else {
    if (direction == 0)
      tk::launch_restrict<float>(F, C, (const float*)in, (float*)out, s, nullptr);
    else
      tk::launch_prolong_add<float>(C, F, (const float*)in, (float*)out, s, nullptr);
  }
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_plan_tail_level(const tk_plan* plan, int* level) {
  if (!plan || !plan->p || !level) {
    set_error("null plan or level");
    return 1;
  }
  *level = plan->p->tail_level;
  return 0;
}

int tk_tail_apply(tk_plan* plan, int prec, const void* b, void* x, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!b || !x) {
    set_error("b and x are required");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  if (P.last_penal < 0.0 || prec != P.last_prec) {
    set_error("tk_tail_apply needs a previous tk_solve or tk_prepare in this precision");
    return 1;
  }
  return tk::apply_tail(P, prec, b, x, (hipStream_t)stream);
}

int tk_plan_persistent_levels(const tk_plan* plan, int* first, int* count) {
  if (!plan || !plan->p || !first || !count) {
    set_error("null plan or output");
    return 1;
  }
  *first = plan->p->pc_first;
  *count = plan->p->pc_nlev;
  return 0;
}

int tk_plan_set_barrier_timeout(tk_plan* plan, double seconds) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  if (!(seconds >= 0.0 && seconds <= 10.0)) {
    set_error("the barrier timeout must be in [0, 10] seconds");
    return 1;
  }
  // the kernels read the realtime clock, which counts at 100 MHz
  plan->p->pc_spin = (unsigned long long)(seconds * 1e8);
  return 0;
}

int tk_coarse_cycle_apply(tk_plan* plan, int prec, const void* b, void* x, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!b || !x) {
    set_error("b and x are required");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  if (P.last_penal < 0.0 || prec != P.last_prec) {
    set_error("tk_coarse_cycle_apply needs a previous tk_solve or tk_prepare in this precision");
    return 1;
  }
  return tk::apply_coarse_cycle(P, prec, b, x, (hipStream_t)stream);
}

int tk_plan_galerkin_info(const tk_plan* plan, int* level, double* bound) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  const tk::Plan& P = *plan->p;
  if (level) *level = P.gal_level;
  if (bound) *bound = P.gal_level >= 0 ? P.lam_gal : P.lam_bound;
  return 0;
}

int tk_smoother_step(tk_plan* plan, int prec, const void* b, const void* x, void* xo, void* d,
                     double c1, double c2, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!b || !x || !xo || !d) {
    set_error("b, x, xo and d are required");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  if (P.last_penal < 0.0 || prec != P.last_prec) {
    set_error("tk_smoother_step needs a previous tk_solve in this precision");
    return 1;
  }
  return tk::smooth_step_fine(P, prec, b, x, xo, d, c1, c2, (hipStream_t)stream);
}

int tk_diag_K(tk_plan* plan, const double* rho, double penal, double* diag, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!rho || !diag) {
    set_error("rho and diag are required");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  tk::launch_stiffness_scale<double>(P, rho, penal, P.lv[0].f64.s, s);
  tk::launch_diag_inv<double>(P, P.lv[0], P.lv[0].f64.s, (double*)nullptr, diag, s);
  P.last_penal = -1.0;
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_prepare(tk_plan* plan, int prec, const double* rho, double penal, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!rho) {
    set_error("rho is required");
    return 1;
  }
  if (!(std::isfinite(penal) && penal > 0)) {
    set_error("penal must be positive, got " + std::to_string(penal));
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  return tk::setup_hierarchy(P, rho, penal, prec, (hipStream_t)stream);
}

int tk_solve(tk_plan* plan, int prec, const double* rho, double penal, const double* F,
             double* U, double rtol, int maxit, int flags, int* iters_out, double* relres_out,
             void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!rho || !F || !U) {
    set_error("rho, F and U are required");
    return 1;
  }
  if (!(std::isfinite(penal) && penal > 0)) {
    set_error("penal must be positive, got " + std::to_string(penal));
    return 1;
  }
  if (!(rtol > 0.0) || maxit < 1) {
    set_error("rtol must be positive and maxit at least 1");
    return 1;
  }
  if (prec != TK_FP64 && prec != TK_MIXED_FP32) {
    set_error("unknown precision");
    return 1;
  }
  // each set-up builds the state of its own precision only (a mixed set-up
  // forms no FP64 inverse diagonals)
  if ((flags & TK_SOLVE_REUSE) && (P.last_penal != penal || prec != P.last_prec ||
                                    rho != P.last_rho || P.setup_mask_gen != P.mask_gen)) {
    set_error("TK_SOLVE_REUSE needs a previous solve or tk_prepare on this plan with the same rho "
              "buffer, penal and precision, and no tk_plan_set_masks since");
    return 1;
  }
  return tk::solve(P, prec, rho, penal, F, U, rtol, maxit, flags, iters_out, relres_out,
                   (hipStream_t)stream);
}

int tk_energy(tk_plan* plan, const double* U, const double* rho, const double* drho, double penal,
              double* ce, double* dc_drho, double* dc_dx, double* dv_dx, double* c_out,
              void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!U || !rho) {
    set_error("U and rho are required");
    return 1;
  }
  if ((dc_dx || dv_dx) && !drho) {
    set_error("drho is required for the chain rule");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool chain = dc_dx != nullptr;
  tk::launch_energy(P, U, rho, drho, penal, ce, dc_drho, chain ? P.ework2 : nullptr,
                    chain ? P.ework3 : nullptr, s);
  if (chain) tk::launch_filter_transpose2(P, P.ework2, dv_dx ? P.ework3 : nullptr, dc_dx, dv_dx, s);
  TK_LAUNCH_CHECK();
  if (c_out) {
    if (tk::read_sc(P, s)) return 2;
    *c_out = P.sc_host->c;
  }
  return 0;
}

int tk_oc_update(tk_plan* plan, const double* x, const double* dc_dx, const double* dv_dx,
                 double move, double target, double beta, double eta, int max_steps, double tol,
                 double* x_new, int* steps_out, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!std::isfinite(move) || move <= 0) {
    set_error("move must be positive, got " + std::to_string(move));
    return 1;
  }
  if (max_steps < 1) {
    set_error("max_bisection_iterations must be at least 1");
    return 1;
  }
  if (tk::check_beta_eta(beta, eta)) return 1;
  if (!x || !dc_dx || !dv_dx || !x_new) {
    set_error("x, dc_dx, dv_dx and x_new are required");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  double tb, den;
  tk::proj_consts(beta, eta, tb, den);
  tk::launch_oc_ratio(P, dc_dx, dv_dx, P.ework0, s);
  tk::launch_oc_normalize(P, P.ework0, s);
  tk::launch_oc_init(P, target, max_steps, tol, 1, s);
  if (tk::oc_bisect(P, x, move, beta, eta, tb, den, max_steps, s)) return 2;
  if (P.sc_host->redo) {
    // a midpoint decided unmeasured might have been the kept one: measure all
    tk::launch_oc_init(P, target, max_steps, tol, 0, s);
    if (tk::oc_bisect(P, x, move, beta, eta, tb, den, max_steps, s)) return 2;
  }
  tk::launch_oc_candidate(P, x, P.ework0, move, 1, x_new, s);
  TK_LAUNCH_CHECK();
  if (steps_out) *steps_out = P.sc_host->steps;
  return 0;
}

int tk_design_stats(tk_plan* plan, const double* x_old, const double* x_new, const double* rho,
                    void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (!rho) {
    set_error("rho is required");
    return 1;
  }
  tk::launch_design_stats(P, x_old, x_new, rho, (hipStream_t)stream);
  TK_LAUNCH_CHECK();
  return 0;
}

int tk_volume_bounds(tk_plan* plan, double beta, double eta, double* vmin, double* vmax,
                     void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (tk::check_beta_eta(beta, eta)) return 1;
  hipStream_t s = (hipStream_t)stream;
  double tb, den;
  tk::proj_consts(beta, eta, tb, den);
  tk::launch_bounds_design(P, 0, P.ework0, s);
  tk::launch_free_sum_projected(P, P.ework0, beta, eta, tb, den, 0, s);
  tk::launch_bounds_design(P, 1, P.ework1, s);
  tk::launch_free_sum_projected(P, P.ework1, beta, eta, tb, den, 1, s);
  TK_LAUNCH_CHECK();
  if (tk::read_sc(P, s)) return 2;
  const double nf = (double)(P.nfree > 0 ? P.nfree : 1);
  if (vmin) *vmin = P.sc_host->sum0 / nf;
  if (vmax) *vmax = P.sc_host->sum1 / nf;
  return 0;
}

int tk_read_results(tk_plan* plan, double* out, int n, void* stream) {
  TK_PLAN_OR_FAIL(plan);
  if (tk::read_sc(P, (hipStream_t)stream)) return 2;
  const tk::Scalars& S = *P.sc_host;
  const double v[10] = {S.c, S.change, S.vol_sum, S.gray_cnt, (double)S.steps, S.best_lam,
                        S.V, S.ratio_scale, (double)P.nfree, (double)S.evals};
  for (int i = 0; i < n && i < 10; ++i) out[i] = v[i];
  return 0;
}

int tk_plan_counters(const tk_plan* plan, long long* out, int n) {
  if (!plan || !plan->p) {
    set_error("null plan");
    return 1;
  }
  const tk::Counters& c = plan->p->cnt;
  const long long v[7] = {c.applies64, c.applies32, c.cg_iters, c.ir_steps,
                          c.solves, c.oc_steps, c.filter_passes};
  for (int i = 0; i < n && i < 7; ++i) out[i] = v[i];
  return 0;
}

const char* tk_last_error(void) { return tk::last_error(); }

}  // extern "C"
