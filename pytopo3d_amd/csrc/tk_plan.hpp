// Plan object: everything one problem size needs on the device, allocated
// once at creation (stencil table, Hs, masks, multigrid hierarchy, Krylov
// vectors, reduction scratch).  Calls on a plan allocate nothing.
#pragma once

#include "tk_common.hpp"

#include <string>
#include <vector>

namespace tk {

template <typename T>
struct LevelBuf {
  T* s = nullptr;      // element stiffness scale, s = Emin + (1 - Emin) rho^p (fine)
  T* dinv = nullptr;   // inverse diagonal (1 on fixed DOFs)
  T* x = nullptr;      // iterate
  T* xa = nullptr;     // ping-pong partner of x
  T* b = nullptr;      // right-hand side (levels >= 1)
  T* r = nullptr;      // residual
  T* d = nullptr;      // Chebyshev direction
};

struct Level {
  Grid g{};
  uint8_t* fixed = nullptr;      // per node, bit c set = DOF c fixed
  LevelBuf<double> f64;
  LevelBuf<float> f32;
};

// Coarse levels (the third and coarser) with more elements than this are
// smoothed with Options::mid_degree / mid_cheb_lo, the smaller ones (those the
// persistent coarse cycle and the tail can hold) with coarse_degree /
// coarse_cheb_lo.
constexpr int kMidMinEl = 32768;

struct Options {
  int smooth_degree = 2;         // Chebyshev degree per pre/post smoothing, two finest levels
  int coarse_degree = 12;        // ... on the coarse levels of <= kMidMinEl elements
  double coarse_cheb_lo = 0.0005; // cheb_lo on those levels
  int mid_degree = 8;            // ... on the coarse levels of more elements
  double mid_cheb_lo = 0.002;    // cheb_lo on those levels
  double cheb_lo = 0.06;         // smoothing interval [cheb_lo, cheb_hi] * lambda_max bound
  double cheb_hi = 1.0;
  int max_levels = kMaxLevels;
  int coarse_max_dof = kCoarseMaxDof;
  int check_every = 8;           // Krylov convergence is read back every N iterations
  int operator_kind = 0;         // 0: element/sector kernel when KE allows it, 1: dense node
                                 // kernel, 2: element kernel with BF16 MFMA FP32 sweeps
  int no_graphs = 0;             // 1: no captured solver chunks
  int filter_kind = 0;           // 0: LDS tiles when the radius allows, 1: stencil gather
  int direction_slots = 0;       // mixed-solver direction ring (0: by memory, else 2..64)
  int galerkin = 0;              // Galerkin operator on the third level: 0 default (off), 1 on, 2 off
  int persistent_coarse = 1;     // 1: the small levels above the tail in one launch per direction
};

struct Counters {
  long long applies64 = 0, applies32 = 0;      // fine-level operator applications
  long long cg_iters = 0, ir_steps = 0, solves = 0;
  long long oc_steps = 0, filter_passes = 0;
};

// direction ring of the mixed solver: at most kRingMax slots, about
// kRingBytes in all
constexpr int kRingMax = 64;
constexpr long long kRingBytes = 2LL << 30;

struct Plan {
  Grid g{};
  double KE[576];                 // host copy, row-major
  double* KE64 = nullptr;         // device
  float* KE32 = nullptr;
  double lam_bound = 0.0;         // lambda_max(diag(KE)^-1 KE): bound of lambda_max(D^-1 K)
  bool sector_ok = false;         // KE block-diagonalises in the Walsh-Hadamard basis
  bool use_el = false;            // element kernel in use
  bool mfma = false;              // FP32 element sweeps in the BF16 MFMA form (operator_kind 2)
  void* kemf = nullptr;           // its KE' operand, BF16, per-lane layout (2 KiB)
  double sec64[45];               // non-zero sector entries of W^T KE W / 64
  float sec32[45];
  double kd[3] = {0, 0, 0};       // KE diagonal per component
  bool kd_ok = false;             // ... equal at all 8 corners (diagonal formed on the fly)
  double* sec64_dev = nullptr;    // device copies (staged into LDS by the element kernel)
  float* sec32_dev = nullptr;
  int ncu = 0;                    // compute units of the plan's device
  // density filter
  double rmin = 0.0;
  int nsten = 0, R = 0;
  int reach = 0;                  // largest per-axis stencil offset
  int* sten_off = nullptr;        // packed (dy+16) | (dx+16)<<8 | (dz+16)<<16
  double* sten_w = nullptr;
  double* Hs = nullptr;
  void* fcols = nullptr;          // column table of the LDS-tiled filter
  int fhalo = 0;                  // its halo (largest stencil offset)
  size_t flds = 0;                // its dynamic LDS bytes
  bool ftiled = false;            // tiled filter available (halo <= 3)
  uint8_t* eflags = nullptr;      // per element, kSolid | kVoid
  long long nfree = 0;
  bool masks_set = false;
  // multigrid hierarchy
  int nlev = 0;
  Level lv[kMaxLevels];
  int ncoarse = 0;                // coarsest DOF count
  double* coarse_inv64 = nullptr; // dense inverse of the coarsest operator (ncoarse^2)
  float* coarse_inv32 = nullptr;
  double* coarse_L = nullptr;     // packed Cholesky factor scratch
  // Krylov vectors (fine level)
  double *r64 = nullptr, *p64 = nullptr, *q64 = nullptr, *x64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
  // flexible-CG directions of the last ring_m iterations (FP32, ndof apart) and
  // their step lengths: the iterate is updated from them every ring_m
  // iterations and at the end instead of in every update pass
  float* pring = nullptr;
  int ring_m = 0;
  bool ring_tried = false;        // allocated by the first mixed solve (ensure_ring)
  double* alpha_hist = nullptr;
  // element work arrays
  double* ework0 = nullptr;       // OC ratio / design fills
  double* ework1 = nullptr;       // OC candidate
  double* ework2 = nullptr;       // chain-rule weights (compliance)
  double* ework3 = nullptr;       // chain-rule weights (volume)
  // reductions
  double* partials = nullptr;
  int partials_len = 0;
  Scalars* sc = nullptr;          // device scalars
  Scalars* sc_host = nullptr;     // pinned mirror
  Scalars* sc_host_dev = nullptr; // its device address (the captured chunks write it)
  Options opt;
  Counters cnt;
  // bookkeeping of the last set-up (TK_SOLVE_REUSE must match all of it)
  double last_penal = -1.0;
  int last_prec = 0;
  const double* last_rho = nullptr;
  long long mask_gen = 0;         // bumped by every tk_plan_set_masks
  long long setup_mask_gen = -1;  // mask_gen at the last set-up
  // captured Krylov chunks, [precision][jacobi]: every launch argument of a
  // chunk is plan-owned (the iterate lives in x64), so a chunk is captured
  // once per plan (on gstream) and replayed on the caller's stream.
  hipStream_t gstream = nullptr;
  // the coarsest factorisation runs on cstream beside the rest of the set-up;
  // the first coarsest solve after it waits for cev_out (coarse_pending)
  hipStream_t cstream = nullptr;
  hipEvent_t cev_in = nullptr, cev_out = nullptr;
  bool coarse_pending = false;
  static constexpr int kChunkSizes = 5;   // chunk lengths 2 * base, base, base / 2, 2, 1
  // [precision][jacobi][length][parity of the chunk's first iteration]
  hipGraphExec_t gexec[2][2][kChunkSizes][2] = {};
  Counters gdelta[2][2][kChunkSizes][2];   // host counters added by one replay
  int lds_limit = 0;              // per-workgroup LDS bytes of the device
  int tail_level = -1;            // level whose V-cycle part runs in one workgroup (-1: none)
  // persistent coarse cycle (tk_pcycle.hip): levels pc_first .. tail_level - 1
  int pc_first = -1;
  int pc_nlev = 0;
  int pc_nb[kMaxLevels][3] = {};  // brick counts along y, x, z per level
  unsigned* pc_sync = nullptr;    // its tickets and barrier flags (kPcSyncWords)
  int pc_box_max = 0, pc_epad_max = 0;   // its LDS layout
  unsigned long long pc_spin = 5000000;  // its barrier wait limit (100 MHz clock ticks: 50 ms)
  bool tail_inv_lds64 = false;    // the tail keeps a copy of the coarsest inverse in LDS
  bool tail_inv_lds32 = false;
  // Galerkin operator of one level (tk_galerkin.hip): per element the Walsh
  // coefficients of its 8 children's scales, the 279-entry coupling table
  int gal_level = -1;             // the level (-1: every level rediscretised)
  bool launch_refused = false;    // a sweep was not launched (missing operand): the solve fails
  double lam_gal = 0.0;           // bound of lambda_max(D^-1 A) on that level
  double* gal_tab64 = nullptr;    // coupling table (kGalN entries)
  float* gal_tab32 = nullptr;
  double* gal_dw = nullptr;       // Walsh-domain corner diagonals (8 x 8 x 3)
  double* gal_sh64 = nullptr;     // 8 coefficients per element of the level
  float* gal_sh32 = nullptr;
};

// Chebyshev coefficients of one smoothing pass on the single-workgroup tail
// (step 0: d = c0 D^-1 r; step k >= 1: d = c1[k] d + c2[k] D^-1 r)
constexpr int kTailMaxDeg = 12;
constexpr int kTailMaxNodes = 512;
struct TailCoef {
  int deg;
  double c0;
  double c1[kTailMaxDeg];
  double c2[kTailMaxDeg];
};

// ---- two-phase reductions (tk_reduce.hip) -----------------------------------
enum FinOp {
  kFinSum0 = 0, kFinSum1, kFinNorm0, kFinNorm1, kFinC,
  kFinAlpha, kFinNorms, kFinUpdate, kFinRz,
  kFinRatioMax, kFinOcVolume, kFinStats, kFinUpdateFlex, kFinRzPR
};
constexpr int kRedBlocks = 2048;   // grid-stride reduction kernels use at most this many blocks
// partials: nullptr = the plan's reduction scratch
void finalize(Plan& P, int op, int G, hipStream_t s, double a0 = 0.0, int i0 = 0,
              const double* partials = nullptr);
inline int red_blocks(long long n) {
  const int b = nblocks(n);
  return b < kRedBlocks ? (b > 0 ? b : 1) : kRedBlocks;
}

// ---- launchers (defined in the .hip files) --------------------------------
// filter
void launch_filter_hs(Plan& P, hipStream_t s);
void launch_filter_project(Plan& P, const double* x, double beta, double eta, double tb_eta,
                           double den, double* rf, double* rho, double* drho, hipStream_t s);
void launch_filter_transpose2(Plan& P, const double* wc, const double* wv, double* oc,
                              double* ov, hipStream_t s);
void launch_free_sum_projected(Plan& P, const double* x, double beta, double eta, double tb_eta,
                               double den, int slot, hipStream_t s);
void launch_bounds_design(Plan& P, int upper, double* out, hipStream_t s);
void launch_count_free(Plan& P, hipStream_t s);
// operator / multigrid
template <typename T>
void launch_stiffness_scale(Plan& P, const double* rho, double penal, T* s_out, hipStream_t s);
void launch_stiffness_scale2(Plan& P, const double* rho, double penal, double* s64, float* s32,
                             hipStream_t s);
template <typename T>
void launch_coarsen_scale(const Level& f, const Level& c, const T* sf, T* sc, hipStream_t s);
void launch_coarsen_fixed(const Level& f, const Level& c, hipStream_t s);
// the element scales of every coarse level, FP64 and FP32 chains together
void launch_coarsen_scale2(Plan& P, hipStream_t s);
// the inverse diagonals of every level (precision T) in one launch
template <typename T>
void launch_diag_inv_levels(Plan& P, hipStream_t s);
template <typename T>
void launch_diag_inv(Plan& P, const Level& L, const T* sl, T* dinv, T* diag_out, hipStream_t s);
template <typename T>
void launch_apply(Plan& P, const Level& L, const T* sl, const T* u, T* y, double* dot_out_slot,
                  int with_dot, hipStream_t s);
template <typename T>
void launch_residual(Plan& P, const Level& L, const T* sl, const T* b, const T* x, T* r,
                     hipStream_t s, const int* done);
// write_d = 0 (the last step of a smoothing pass) lets the element kernel skip
// storing the direction d, which nothing reads afterwards.
template <typename T>
void launch_cheb_first(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b,
                       const T* x, T* xo, T* d, double c0, int x_zero, int write_d, hipStream_t s);
// dot_blocks != nullptr asks the step for the partials of b.xo (one per
// workgroup, in the plan's reduction scratch); it receives their count, or 0
// when this launch cannot form them (the caller then reduces separately).
template <typename T>
void launch_cheb_step(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b,
                      const T* x, T* xo, T* d, double c1, double c2, int write_d, hipStream_t s,
                      int* dot_blocks = nullptr);
template <typename T>
void launch_restrict(const Level& f, const Level& c, const T* rf, T* bc, hipStream_t s,
                     const int* done);
template <typename T>
void launch_prolong_add(const Level& c, const Level& f, const T* xc, T* xf, hipStream_t s,
                        const int* done);
hipError_t coarse_setup_attr(size_t lds);
bool operator_setup(Plan& P);
bool filter_tiled_setup(Plan& P, const std::vector<int>& off, const std::vector<double>& w);
void tiled_filter_project(Plan& P, const double* x, double beta, double eta, double tb, double den,
                          double* rf, double* rho, double* drho, hipStream_t s);
void tiled_filter_plain(Plan& P, const double* v, double* out, hipStream_t s);
void tiled_free_sum(Plan& P, const double* x, double beta, double eta, double tb, double den,
                    int slot, hipStream_t s);
void tiled_oc_volume(Plan& P, const double* cand, double beta, double eta, double tb, double den,
                     hipStream_t s);
template <typename T>
void el_apply(Plan& P, const Level& L, const T* sl, const T* u, T* y, int with_dot,
              const int* done, hipStream_t s);
template <typename T>
void el_residual(Plan& P, const Level& L, const T* sl, const T* b, const T* x, T* r,
                 const int* done, hipStream_t s);
template <typename T>
void el_cheb(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, const T* x, T* xo,
             T* d, double c0, double c1, double c2, int first, int write_d, const int* done,
             hipStream_t s, int* dot_blocks = nullptr);
template <typename T>
void el_cheb_zero2(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, T* xo, T* d,
                   double c0, double c1, double c2, int write_d, const int* done, hipStream_t s);
void el_apply_dir(Plan& P, const float* z32, const float* p32o, float* p32n, double* y,
                  hipStream_t s);
int el_apply_dir_ring(Plan& P, const float* z32, int have_old, double* y, hipStream_t s);
template <typename T>
void el_cheb_prolong(Plan& P, const Level& L, const Level& C, const T* xc, const T* sl,
                     const T* dinv, const T* b, const T* x, T* xo, T* d, double c0, int write_d,
                     const int* done, hipStream_t s);
void launch_coarse_setup(Plan& P, hipStream_t s);
// Galerkin level (tk_galerkin.hip)
double scaled_lambda_max(const double* M);
bool galerkin_setup(Plan& P);
template <typename T>
void launch_gal_coeffs(Plan& P, hipStream_t s);
template <typename T>
void launch_gal_diag(Plan& P, hipStream_t s);
// the second-coarsest level's V-cycle part and the coarsest solve in one
// workgroup (tk_tail.hip); tail_setup picks the level when it fits
void tail_setup(Plan& P);
template <typename T>
void launch_tail(Plan& P, const T* b, T* x, const TailCoef& cf, hipStream_t s, const int* done);
// persistent coarse cycle (tk_pcycle.hip); pcycle_setup picks its levels
constexpr int kPcMaxLev = 4;
constexpr int kPcSyncWords = 64 + 2 * 32 * 32;   // tickets, then 2 x 32 flag lines
void pcycle_setup(Plan& P);
template <typename T>
void launch_pcycle_down(Plan& P, const TailCoef& cf, hipStream_t s);
template <typename T>
void launch_pcycle_up(Plan& P, const TailCoef& cf, hipStream_t s);
template <typename T>
void launch_coarse_solve(Plan& P, const T* b, T* x, hipStream_t s);
// energy
void launch_energy(Plan& P, const double* U, const double* rho, const double* drho, double penal,
                   double* ce, double* dc_drho, double* wc, double* wv, hipStream_t s);
// optimality criteria
void launch_oc_ratio(Plan& P, const double* dc, const double* dv, double* ratio, hipStream_t s);
void launch_oc_normalize(Plan& P, double* ratio, hipStream_t s);
void launch_oc_candidate(Plan& P, const double* x, const double* ratio, double move, int use_best,
                         double* cand, hipStream_t s);
void launch_oc_step(Plan& P, const double* x, const double* ratio, double move, double* cand,
                    double beta, double eta, double tb_eta, double den, hipStream_t s);
void launch_design_stats(Plan& P, const double* x_old, const double* x_new, const double* rho,
                         hipStream_t s);

// ---- host-side drivers (tk_solver.hip) ------------------------------------
int solve(Plan& P, int prec, const double* rho, double penal, const double* F, double* U,
          double rtol, int maxit, int flags, int* iters_out, double* relres_out, hipStream_t s);
int setup_hierarchy(Plan& P, const double* rho, double penal, int prec, hipStream_t s);
int apply_level(Plan& P, int prec, int l, const void* u, void* y, hipStream_t s);
int apply_tail(Plan& P, int prec, const void* b, void* x, hipStream_t s);
int apply_coarse_cycle(Plan& P, int prec, const void* b, void* x, hipStream_t s);

// ---- errors -----------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

}  // namespace tk

#define TK_HIP(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::tk::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));     \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)
