// Linear solve K U = F on the free DOFs: preconditioned conjugate gradients
// with a geometric-multigrid V-cycle (or Jacobi), in FP64, or as FP64 flexible
// conjugate gradients around an FP32 multigrid (or Jacobi) correction.
//
// Replaces the solver factory of pytopo3d/utils/solver.py:18-73 (PyPardiso /
// SuperLU direct solve, CuPy CG+Jacobi) at its call site
// pytopo3d/core/optimizer.py:591-594.  All step lengths, dot products and the
// refinement residual are FP64; the host only reads the device scalars every
// Options::check_every iterations.
#include "tk_plan.hpp"

#include <cmath>
#include <cstdio>
#include <utility>

namespace tk {

// ------------------------------------------------------------------ vector kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int n, T* __restrict__ a, T v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_zero(int n, const T* __restrict__ b,
                                                       T* __restrict__ r, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = b[i];
  x[i] = T(0);
}

// b.b and r.r partials (kFinNorms arms the iteration)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_norms(int n, const T* __restrict__ b,
                                                      const T* __restrict__ r, double* partials) {
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = b[i], ri = r[i];
    v[0] += bi * bi;
    v[1] += ri * ri;
  }
  block_partials<2>(v, partials);
}

// x += alpha p; r -= alpha q; r.r partials (kFinUpdate tests convergence)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int n, T* __restrict__ x, T* __restrict__ r,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const int* __restrict__ done,
                                                       double* partials, const Scalars* sc) {
  if (*done) return;
  const T alpha = (T)sc->alpha;
  double v[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
  }
  block_partials<1>(v, partials);
}

// r.z partials (kFinRz computes beta)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_rz(int n, const T* __restrict__ r,
                                                   const T* __restrict__ z,
                                                   const int* __restrict__ done, double* partials) {
  if (*done) return;
  double v[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[0] += (double)r[i] * (double)z[i];
  block_partials<1>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dir(int n, const T* __restrict__ z,
                                                    T* __restrict__ p, int first,
                                                    const int* __restrict__ done,
                                                    const Scalars* __restrict__ sc) {
  if (*done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (first) p[i] = z[i];
  else p[i] = z[i] + (T)sc->beta * p[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const T* __restrict__ dinv,
                                                    const T* __restrict__ r, T* __restrict__ z,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = dinv[i] * r[i];
}

// The reduced system's right-hand side out = F and, when U is given, the
// initial iterate x = U, both 0 on fixed DOFs.
__global__ __launch_bounds__(kBlock) void k_mask_rhs(int nnode, const uint8_t* __restrict__ fixed,
                                                      const double* __restrict__ F,
                                                      double* __restrict__ out,
                                                      const double* __restrict__ U,
                                                      double* __restrict__ x) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * n + c] = ((f >> c) & 1) ? 0.0 : F[3 * n + c];
  if (U) {
#pragma unroll
    for (int c = 0; c < 3; ++c) x[3 * n + c] = ((f >> c) & 1) ? 0.0 : U[3 * n + c];
  }
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double* __restrict__ a,
                                                    float* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---- mixed precision: FP64 flexible CG around an FP32 multigrid correction ----
// The FP32 preconditioner sees r / zs and its output is scaled back by zs, with
// zs a recent residual norm (sc->zs), which keeps the FP32 values near unit
// size whatever the load magnitude.
// r32 = r64 / zs (first correction: zs = ||r0||)
__global__ __launch_bounds__(kBlock) void k_to_f32_unit(int n, const double* __restrict__ r64,
                                                         float* __restrict__ r32,
                                                         const int* __restrict__ done,
                                                         const Scalars* __restrict__ sc) {
  if (*done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double zs = sc->zs;
  const double inv = zs > 0.0 ? 1.0 / zs : 0.0;
  r32[i] = (float)(r64[i] * inv);
}

// x += alpha p (p stored in FP32); r -= alpha q; r32 = r / ||r_old|| (the norm
// before this update; kFinUpdateFlex records it as zs); partials of r.r and of
// q.z with z = zs z32 the correction p was formed from (the next beta).
// With a direction ring (ring != nullptr; p32 unused) the iterate is not
// touched here: alpha goes to alpha_hist[iter % m], and the pass whose slot is
// m - 1 adds all m ring directions to x, oldest first -- the same sums, in the
// same order, as one update per iteration.
// G_pq > 0: alpha = rz / p.Kp is formed here from the G_pq partials of p.Kp
// (every block sums them in the same order, so all hold the same alpha; block
// 0 stores it), in place of a separate finalize launch; the pass's own
// partials then go to `partials`, which must not overlap pq_partials.
__global__ __launch_bounds__(kBlock) void k_cg_update_flex(int n, double* __restrict__ x,
                                                            double* __restrict__ r,
                                                            const float* __restrict__ p32,
                                                            const double* __restrict__ q,
                                                            float* __restrict__ r32,
                                                            const float* __restrict__ z32,
                                                            const int* __restrict__ done,
                                                            double* partials, Scalars* sc,
                                                            const float* __restrict__ ring, int m,
                                                            double* __restrict__ alpha_hist,
                                                            const double* __restrict__ pq_partials,
                                                            int G_pq) {
  if (*done) return;
  double alpha;
  if (G_pq > 0) {
    __shared__ double s_pq;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < G_pq; b += blockDim.x) v[0] += pq_partials[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) s_pq = v[0];
    __syncthreads();
    const double pq = s_pq;
    const bool ok = pq > 0.0;   // else breakdown (or NaN): stop the iteration
    alpha = ok ? sc->rz / pq : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pq = pq;
      sc->alpha = alpha;
      if (!ok) {
        sc->done = 1;
        sc->stalled = 1;
      }
    }
    if (!ok) return;
  } else {
    alpha = sc->alpha;
  }
  const double rr = sc->rr;
  const double zs = sc->zs;
  const double inv = rr > 0.0 ? 1.0 / sqrt(rr) : 0.0;
  int slot = 0;
  if (ring) {
    slot = sc->iter % m;
    if (blockIdx.x == 0 && threadIdx.x == 0) alpha_hist[slot] = alpha;
  }
  const bool fold = ring && slot == m - 1;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!ring) {
      x[i] += alpha * (double)p32[i];
    } else if (fold) {
      double xi = x[i];
      for (int j = 0; j < m - 1; ++j) xi += alpha_hist[j] * (double)ring[(long long)j * n + i];
      x[i] = xi + alpha * (double)ring[(long long)(m - 1) * n + i];
    }
    const double qi = q[i];
    const double ri = r[i] - alpha * qi;
    r[i] = ri;
    r32[i] = (float)(ri * inv);
    v[0] += ri * ri;
    v[1] += qi * (zs * (double)z32[i]);
  }
  block_partials<2>(v, partials);
}

// U = x + sum_{j < c} alpha_hist[j] ring_j, in slot order: the ring
// directions not yet added to the iterate when the solve stopped.
__global__ __launch_bounds__(kBlock) void k_ring_fold_out(int n, const double* __restrict__ x,
                                                           const float* __restrict__ ring, int c,
                                                           const double* __restrict__ alpha_hist,
                                                           double* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi = x[i];
  // unrolled so that the loads of several slots are in flight together (the
  // sum keeps the slot order)
#pragma unroll 8
  for (int j = 0; j < c; ++j) xi += alpha_hist[j] * (double)ring[(long long)j * n + i];
  U[i] = xi;
}

// z = zs z32; partials of r.z and (unless first) q.z, q = K p_old
__global__ __launch_bounds__(kBlock) void k_cg_rz_flex(int n, const double* __restrict__ r,
                                                        const double* __restrict__ q,
                                                        const float* __restrict__ z32,
                                                        int first, const int* __restrict__ done,
                                                        double* partials, const Scalars* sc) {
  if (*done) return;
  const double nrm = sc->zs;
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double z = nrm * (double)z32[i];
    v[0] += r[i] * z;
    if (!first) v[1] += q[i] * z;
  }
  block_partials<2>(v, partials);
}

// p32n = fp32(zs z32 + beta p32o) and its FP64 copy (operators without the
// fused direction update); p32o == nullptr: beta = 0
__global__ __launch_bounds__(kBlock) void k_cg_dir_flex(int n, const float* __restrict__ z32,
                                                         const float* __restrict__ p32o,
                                                         float* __restrict__ p32n,
                                                         double* __restrict__ p64,
                                                         const int* __restrict__ done,
                                                         const Scalars* __restrict__ sc) {
  if (*done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // formed in FP32 with the fused operator sweep's rounding order
  const float pb = p32o ? (float)sc->beta * p32o[i] : 0.0f;
  const float pf = __builtin_fmaf((float)sc->zs, z32[i], pb);
  p32n[i] = pf;
  p64[i] = (double)pf;
}

// ------------------------------------------------------------------ helpers
template <typename T> static LevelBuf<T>& buf(Level& L);
template <> LevelBuf<double>& buf<double>(Level& L) { return L.f64; }
template <> LevelBuf<float>& buf<float>(Level& L) { return L.f32; }

// The solver scalars copied to the host mirror by a kernel (one captured as
// the last node of every Krylov chunk, so the host only waits for the chunk).
__global__ __launch_bounds__(64) void k_scalars_out(const Scalars* __restrict__ sc,
                                                     Scalars* __restrict__ host) {
  constexpr int n = (int)(sizeof(Scalars) / sizeof(int));
  static_assert(sizeof(Scalars) % sizeof(int) == 0, "whole words");
  const int* a = reinterpret_cast<const int*>(sc);
  int* b = reinterpret_cast<int*>(host);
  for (int i = threadIdx.x; i < n; i += 64) b[i] = a[i];
}

static int read_scalars(Plan& P, hipStream_t s) {
  TK_HIP(hipMemcpyAsync(P.sc_host, P.sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  TK_HIP(hipStreamSynchronize(s));
  return 0;
}

// Chebyshev degree and interval low end (fraction of the lambda_max bound) of
// level l.  The coarse levels are latency-bound: a longer smoother over a
// wider interval costs little there and strengthens the cycle; the smallest
// ones (a few microseconds per step in the persistent cycle and the tail) take
// the longest.
static void level_cheb(const Plan& P, int l, int& deg, double& lo) {
  if (l < 2) {
    deg = P.opt.smooth_degree;
    lo = P.opt.cheb_lo;
  } else if (P.lv[l].g.nele > kMidMinEl) {
    deg = P.opt.mid_degree;
    lo = P.opt.mid_cheb_lo;
  } else {
    deg = P.opt.coarse_degree;
    lo = P.opt.coarse_cheb_lo;
  }
}

// Chebyshev smoothing of level l: `deg` steps from x (x_zero: from 0).
// Returns the buffer holding the result (x or its ping-pong partner).
constexpr int kProlongFuseMaxNodes = 100000;

// xc != nullptr (post-smoothing, element kernel): the smoothing starts from
// x + P xc, the coarse-grid correction being prolongated inside the first step.
// dot_blocks != nullptr: the last step also leaves the partials of b.x (see
// launch_cheb_step); their count, 0 if it could not.
template <typename T>
static T* smooth(Plan& P, int l, const T* b, T* x, T* xa, int x_zero, hipStream_t s,
                 const T* xc = nullptr, int* dot_blocks = nullptr) {
  Level& L = P.lv[l];
  int deg;
  double lo;
  level_cheb(P, l, deg, lo);
  LevelBuf<T>& B = buf<T>(L);
  // the Galerkin level has its own bound of lambda_max(D^-1 A)
  const double lam = l == P.gal_level ? P.lam_gal : P.lam_bound;
  const double lmax = lam * P.opt.cheb_hi;
  const double lmin = lam * lo;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  int k = 1;
  if (x_zero && P.use_el && deg >= 2) {
    // steps 1 and 2 from x = 0 in one sweep
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    if (&L == &P.lv[0]) {
      if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
    }
    el_cheb_zero2<T>(P, L, B.s, B.dinv, b, xa, B.d, 1.0 / theta, rho_new * rho_old,
                     2.0 * rho_new / delta, deg > 2, &P.sc->done, s);
    std::swap(x, xa);
    rho_old = rho_new;
    k = 2;
  } else if (xc) {
    if (&L == &P.lv[0]) {
      if (sizeof(T) == 8) P.cnt.applies64++; else P.cnt.applies32++;
    }
    el_cheb_prolong<T>(P, L, P.lv[l + 1], xc, B.s, B.dinv, b, x, xa, B.d, 1.0 / theta,
                       deg > 1, &P.sc->done, s);
    std::swap(x, xa);
  } else {
    launch_cheb_first<T>(P, L, B.s, B.dinv, b, x, xa, B.d, 1.0 / theta, x_zero,
                         deg > 1, s);
    std::swap(x, xa);
  }
  if (dot_blocks) *dot_blocks = 0;
  for (; k < deg; ++k) {
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    const bool last = k + 1 == deg;
    launch_cheb_step<T>(P, L, B.s, B.dinv, b, x, xa, B.d, rho_new * rho_old, 2.0 * rho_new / delta,
                        !last, s, last ? dot_blocks : nullptr);
    std::swap(x, xa);
    rho_old = rho_new;
  }
  return x;
}

// The coefficients smooth() uses on level l (same recurrence), for the tail.
static TailCoef tail_coefs(const Plan& P, int l) {
  TailCoef cf{};
  double lo;
  level_cheb(P, l, cf.deg, lo);
  const double lmax = P.lam_bound * P.opt.cheb_hi;
  const double lmin = P.lam_bound * lo;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  cf.c0 = 1.0 / theta;
  for (int k = 1; k < cf.deg && k < kTailMaxDeg; ++k) {
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    cf.c1[k] = rho_new * rho_old;
    cf.c2[k] = 2.0 * rho_new / delta;
    rho_old = rho_new;
  }
  return cf;
}

// Make s wait for the coarsest factorisation of the last set-up, once.  The
// first coarsest solve after a set-up is always launched eagerly (before any
// Krylov chunk is captured), so no captured graph holds this wait.
static void join_coarse(Plan& P, hipStream_t s) {
  if (!P.coarse_pending) return;
  (void)hipStreamWaitEvent(s, P.cev_out, 0);
  P.coarse_pending = false;
}

// z = V-cycle(b) on the hierarchy; returns the pointer holding z (a level-0
// buffer).  dot_blocks: as for smooth, for the last fine-level step (b.z).
template <typename T>
static T* vcycle(Plan& P, const T* b0, hipStream_t s, int* dot_blocks) {
  const int nl = P.nlev;
  const int* done = &P.sc->done;
  T* cur[kMaxLevels];
  T* alt[kMaxLevels];
  const T* rhs[kMaxLevels];
  rhs[0] = b0;
  // from the tail level down, one workgroup runs the rest of the cycle; the
  // persistent coarse cycle runs the levels above it, from pc_first
  const int lt = P.tail_level;
  const bool pc = lt >= 1 && P.pc_nlev > 0;
  const int descend_to = lt >= 1 ? (pc ? P.pc_first : lt) : nl - 1;
  for (int l = 0; l < descend_to; ++l) {
    LevelBuf<T>& B = buf<T>(P.lv[l]);
    T* x = smooth<T>(P, l, rhs[l], B.x, B.xa, 1, s);
    cur[l] = x;
    alt[l] = (x == B.x) ? B.xa : B.x;
    launch_residual<T>(P, P.lv[l], B.s, rhs[l], x, B.r, s, done);
    LevelBuf<T>& C = buf<T>(P.lv[l + 1]);
    launch_restrict<T>(P.lv[l], P.lv[l + 1], B.r, C.b, s, done);
    rhs[l + 1] = C.b;
  }
  if (lt >= 1) {
    const TailCoef cf = tail_coefs(P, lt);
    if (pc) launch_pcycle_down<T>(P, cf, s);   // ends with the tail level's b
    LevelBuf<T>& C = buf<T>(P.lv[lt]);
    join_coarse(P, s);   // the tail reads the coarsest inverse
    launch_tail<T>(P, pc ? C.b : rhs[lt], C.x, cf, s, &P.sc->done);
    cur[lt] = C.x;
    if (pc) {
      launch_pcycle_up<T>(P, cf, s);   // the first persistent level's x
      cur[P.pc_first] = buf<T>(P.lv[P.pc_first]).x;
    }
  } else {
    LevelBuf<T>& C = buf<T>(P.lv[nl - 1]);
    join_coarse(P, s);
    launch_coarse_solve<T>(P, rhs[nl - 1], C.x, s);
    cur[nl - 1] = C.x;
  }
  for (int l = descend_to - 1; l >= 0; --l) {
    // small levels: the prolongation rides in the first smoothing sweep (one
    // launch less); large levels: the separate sweep is faster than gathering
    // the coarse correction inside the register-heavy smoothing kernel
    int* db = l == 0 ? dot_blocks : nullptr;
    if (P.use_el && P.lv[l].g.nnode <= kProlongFuseMaxNodes) {
      cur[l] = smooth<T>(P, l, rhs[l], cur[l], alt[l], 0, s, cur[l + 1], db);
    } else {
      launch_prolong_add<T>(P.lv[l + 1], P.lv[l], cur[l + 1], cur[l], s, done);
      cur[l] = smooth<T>(P, l, rhs[l], cur[l], alt[l], 0, s, nullptr, db);
    }
  }
  return cur[0];
}

// dot_blocks: receives the count of r.z partials the preconditioner's last
// sweep left (0: none; the caller reduces r.z itself).
template <typename T>
static const T* precondition(Plan& P, const T* r, T* zbuf, int jacobi, hipStream_t s,
                             int* dot_blocks) {
  *dot_blocks = 0;
  if (jacobi || P.nlev == 1) {
    if (P.nlev == 1 && !jacobi) {
      join_coarse(P, s);
      launch_coarse_solve<T>(P, r, zbuf, s);
      return zbuf;
    }
    const int n = P.g.ndof;
    hipLaunchKernelGGL(k_jacobi<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, buf<T>(P.lv[0]).dinv,
                       r, zbuf, &P.sc->done);
    return zbuf;
  }
  return vcycle<T>(P, r, s, dot_blocks);
}

// ------------------------------------------------------------------ chunk runner
static void add_counters(Counters& a, const Counters& d, int sign) {
  a.applies64 += sign * d.applies64;
  a.applies32 += sign * d.applies32;
  a.cg_iters += sign * d.cg_iters;
  a.ir_steps += sign * d.ir_steps;
  a.solves += sign * d.solves;
  a.oc_steps += sign * d.oc_steps;
  a.filter_passes += sign * d.filter_passes;
}

// Runs `body(stream, m, par)` -- m Krylov iterations, the first of parity par,
// whose launches all return at once after the device-side `done` flag is set --
// in chunks until the flag is set, reading the device scalars back after every
// chunk.  Chunk lengths are taken from {2 base, base, base / 2, 2, 1} (base =
// check_every rounded up to an even count): the next length is the largest one
// not above the iterations still needed, predicted from the residual decay
// since the previous read (from the initial residual for the first), so few
// iterations run past convergence (each costs some 65 launches that return at
// once, ~0.5 ms at 256x128x128, against ~50 us for one more read): the
// prediction aims at three quarters of the estimate, since the convergence of
// CG speeds up.  The parity matters to bodies that alternate buffers between
// iterations (par0 >= 0); it starts at par0 and flips after every chunk of odd
// length (par0 < 0: parity-free body, one graph per length).  With graphs
// enabled every (length, parity) is captured on the plan's stream before the
// first chunk runs and replayed from then on (the kernels, their order and
// arguments are the same in every such chunk), so the device runs it without
// host launch gaps.
template <class Body>
static int run_chunks(Plan& P, int prec, int jacobi, int maxit, int par0, Body body,
                      hipStream_t s) {
  bool graphs = !P.opt.no_graphs;
  if (graphs && !P.gstream) {
    if (hipStreamCreateWithFlags(&P.gstream, hipStreamNonBlocking) != hipSuccess) {
      set_error("could not create the capture stream");
      return 2;
    }
  }
  int base = P.opt.check_every > 0 ? P.opt.check_every : 8;
  base += base & 1;
  int sizes[Plan::kChunkSizes] = {2 * base, base, base / 2 + ((base / 2) & 1), 2, 1};
  const int pa = prec ? 1 : 0, ja = jacobi ? 1 : 0;
  // capture a (length, parity) at its first use, so later solves never pay it
  auto capture = [&](int k, int par) -> int {
    const int m = sizes[k];
    hipGraphExec_t& ex = P.gexec[pa][ja][k][par];
    Counters& delta = P.gdelta[pa][ja][k][par];
    if (!ex) {
      const Counters before = P.cnt;
      TK_HIP(hipStreamBeginCapture(P.gstream, hipStreamCaptureModeThreadLocal));
      body(P.gstream, m, par);
      hipLaunchKernelGGL(k_scalars_out, dim3(1), dim3(64), 0, P.gstream, P.sc, P.sc_host_dev);
      hipGraph_t g = nullptr;
      TK_HIP(hipStreamEndCapture(P.gstream, &g));
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        ex = nullptr;
        set_error(std::string("HIP error: ") + hipGetErrorString(e) + " (graph instantiation)");
        return 2;
      }
      delta = P.cnt;
      add_counters(delta, before, -1);
      P.cnt = before;
    }
    return 0;
  };
  const bool parity = par0 >= 0;
  if (graphs)
    for (int k = 0; k < Plan::kChunkSizes; ++k)
      for (int q = 0; q < (parity ? 2 : 1); ++q)
        if (int rc = capture(k, q)) return rc;
  auto run = [&](int k, int par) -> int {
    if (!graphs) {
      body(s, sizes[k], par);
      return 0;
    }
    // the chunk runs on the caller's stream; its last kernel has copied the
    // scalars to the host mirror
    hipGraphExec_t& ex = P.gexec[pa][ja][k][par];
    Counters& delta = P.gdelta[pa][ja][k][par];
    TK_HIP(hipGraphLaunch(ex, s));
    add_counters(P.cnt, delta, 1);
    return 0;
  };
  int need = sizes[1];             // first round: base iterations
  int par = parity ? par0 : 0;
  // the decay is measured from the initial residual for the first prediction
  // (kFinNorms left it in rr0; the iteration count starts at 0)
  int it_prev = 0;
  double rr_prev = -1.0;
  for (long long guard = 0; guard < (long long)maxit + 8; ++guard) {
    // one round of `need` iterations: the captured chunks whose lengths sum
    // to it, launched back to back, then one look at the scalars
    for (int left = need; left > 0;) {
      int k = Plan::kChunkSizes - 1;
      for (int j = 0; j < Plan::kChunkSizes; ++j)
        if (sizes[j] <= left) { k = j; break; }
      if (int rc = run(k, par)) return rc;
      if (parity) par ^= sizes[k] & 1;
      left -= sizes[k];
    }
    if (graphs) {
      TK_HIP(hipStreamSynchronize(s));
    } else if (read_scalars(P, s)) {
      return 2;
    }
    const Scalars& H = *P.sc_host;
    if (H.done) break;
    if (rr_prev < 0.0) rr_prev = H.rr0;
    // iterations still needed at the observed residual decay rate
    const double target = H.tol2 * H.bb;
    need = sizes[0];
    if (it_prev < H.iter && rr_prev > 0.0 && H.rr > 0.0 && H.rr < rr_prev && target > 0.0) {
      const double per_it = std::log(H.rr / rr_prev) / (double)(H.iter - it_prev);
      const double est = 0.75 * std::log(target / H.rr) / per_it;
      if (est < (double)sizes[0]) need = est < 1.0 ? 1 : (int)std::ceil(est);
    }
    it_prev = H.iter;
    rr_prev = H.rr;
  }
  return 0;
}

// PCG on K x = b (fine level, precision T).  x is in/out (x_zero: start at 0).
// Returns 0, or 2 on a HIP error.  Iterations and the final recurrence
// residual are left in P.sc_host.
template <typename T>
static int pcg(Plan& P, const T* b, T* x, T* r, T* p, T* q, int x_zero, double rtol, int maxit,
               int jacobi, hipStream_t s) {
  const int n = P.g.ndof;
  Level& L0 = P.lv[0];
  LevelBuf<T>& B0 = buf<T>(L0);
  const int* done = &P.sc->done;
  if (x_zero) {
    hipLaunchKernelGGL(k_copy_zero<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, b, r, x);
  } else {
    launch_residual<T>(P, L0, B0.s, b, x, r, s, nullptr);
  }
  hipLaunchKernelGGL(k_cg_norms<T>, dim3(red_blocks(n)), dim3(kBlock), 0, s, n, b, r, P.partials);
  finalize(P, kFinNorms, red_blocks(n), s, rtol * rtol, maxit);
  // r.z: from the V-cycle's last fine sweep when it formed the partials
  auto reduce_rz = [&](const T* zz, int G, int first, hipStream_t st) {
    if (G == 0) {
      hipLaunchKernelGGL(k_cg_rz<T>, dim3(red_blocks(n)), dim3(kBlock), 0, st, n, r, zz, done,
                         P.partials);
      G = red_blocks(n);
    }
    finalize(P, kFinRz, G, st, 0.0, first);
  };
  // One iteration: z = M r (into q when the preconditioner needs a buffer: q
  // is read by nothing after the update), r.z and beta, p = z + beta p,
  // q = K p, alpha, x and r, and the convergence test last -- so a chunk
  // ends right after a test, and the solve that converges at a chunk's end
  // runs no preconditioner for nothing.
  auto iteration = [&](int first, hipStream_t st) {
    int Gk = 0;
    const T* zz = precondition<T>(P, r, q, jacobi, st, &Gk);
    reduce_rz(zz, Gk, first, st);
    hipLaunchKernelGGL(k_cg_dir<T>, dim3(nblocks(n)), dim3(kBlock), 0, st, n, zz, p, first, done,
                       P.sc);
    launch_apply<T>(P, L0, B0.s, p, q, nullptr, 1, st);
    hipLaunchKernelGGL(k_cg_update<T>, dim3(red_blocks(n)), dim3(kBlock), 0, st, n, x, r, p, q,
                       done, P.partials, P.sc);
    finalize(P, kFinUpdate, red_blocks(n), st);
  };
  iteration(1, s);
  auto body = [&](hipStream_t st, int m, int) {
    for (int k = 0; k < m; ++k) iteration(0, st);
  };
  int rc = run_chunks(P, 0, jacobi, maxit, -1, body, s);
  if (rc) return rc;
  if (P.launch_refused) return 2;   // tk_last_error() names the sweep
  if (hipGetLastError() != hipSuccess) {
    set_error("kernel launch failed in pcg");
    return 2;
  }
  return 0;
}

// FP64 flexible PCG whose preconditioner is the FP32 V-cycle (or FP32 Jacobi)
// applied to the scaled FP64 residual: every correction is computed in FP32,
// the search direction is stored in FP32, while q = K p, the residual
// recurrence, dot products and step lengths stay FP64, so the iteration
// reaches FP64 accuracy whatever the stiffness contrast.  The direction update
// is the flexible (Polak-Ribiere) one,
//   beta_k = (r_k.z_k - r_k.z_{k-1}) / (r_{k-1}.z_{k-1}),
//   r_k.z_{k-1} = r_{k-1}.z_{k-1} - alpha_{k-1} q_{k-1}.z_{k-1},
// so it needs r.z (left by the V-cycle's last fine sweep, from the FP32
// residual it smooths against) and q.z (from the update pass): no pass of its
// own over the FP64 vectors.
static int pcg_mixed(Plan& P, const double* b, double* x, int x_zero, double rtol, int maxit,
                     int jacobi, hipStream_t s) {
  const int n = P.g.ndof;
  Level& L0 = P.lv[0];
  const int* done = &P.sc->done;
  double *r = P.r64, *q = P.q64, *p64 = P.p64;
  float* pbuf[2] = {P.p32, P.x32};   // direction, double-buffered (read old, write new)
  float* r32 = L0.f32.b;
  if (x_zero) hipLaunchKernelGGL(k_copy_zero<double>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, b, r, x);
  else launch_residual<double>(P, L0, L0.f64.s, b, x, r, s, nullptr);
  hipLaunchKernelGGL(k_cg_norms<double>, dim3(red_blocks(n)), dim3(kBlock), 0, s, n, b, r,
                     P.partials);
  finalize(P, kFinNorms, red_blocks(n), s, rtol * rtol, maxit);
  hipLaunchKernelGGL(k_to_f32_unit, dim3(nblocks(n)), dim3(kBlock), 0, s, n, r, r32, done, P.sc);
  auto correction = [&](int first, hipStream_t st) -> const float* {
    int G = 0;
    const float* z = precondition<float>(P, r32, P.q32, jacobi, st, &G);
    if (G > 0) {
      // partials of r32.z32: r.z = zs^2 r32.z32
      finalize(P, kFinRzPR, G, st, 0.0, first | 2);
    } else {
      hipLaunchKernelGGL(k_cg_rz_flex, dim3(red_blocks(n)), dim3(kBlock), 0, st, n, r, q, z, 1,
                         done, P.partials, P.sc);
      finalize(P, kFinRzPR, red_blocks(n), st, 0.0, first);
    }
    return z;
  };
  const float* z32 = correction(1, s);
  // iteration i writes direction pbuf[i & 1] and reads pbuf[(i - 1) & 1]
  // (none for i = 0); a chunk's kernels and arguments depend on its length
  // and the parity of its first iteration (run_chunks tracks it).
  // with the element operator the directions go to the plan's ring (the
  // iterate is updated from it every ring_m iterations and by the caller at
  // the end, ring_fold_out); else pnew / pold alternate
  const bool ring = P.use_el && P.pring;
  // the update pass reads the p.Kp partials at the start of the scratch while
  // other blocks already write theirs: those go to the upper half
  double* upd_partials = ring ? P.partials + P.partials_len / 2 : P.partials;
  auto iteration = [&](float* pnew, const float* pold, hipStream_t st) {
    int G_pq = 0;
    if (ring) {
      G_pq = el_apply_dir_ring(P, z32, pold != nullptr, q, st);
    } else if (P.use_el) {
      el_apply_dir(P, z32, pold, pnew, q, st);
    } else {
      hipLaunchKernelGGL(k_cg_dir_flex, dim3(nblocks(n)), dim3(kBlock), 0, st, n, z32, pold, pnew,
                         p64, done, P.sc);
      launch_apply<double>(P, L0, L0.f64.s, p64, q, nullptr, 1, st);
    }
    hipLaunchKernelGGL(k_cg_update_flex, dim3(red_blocks(n)), dim3(kBlock), 0, st, n, x, r, pnew,
                       q, r32, z32, done, upd_partials, P.sc, ring ? P.pring : nullptr,
                       P.ring_m, P.alpha_hist, P.partials, G_pq);
    finalize(P, kFinUpdateFlex, red_blocks(n), st, 0.0, 0, upd_partials);
  };
  iteration(pbuf[0], nullptr, s);
  // the correction opens each chunk iteration, so a chunk ends right after a
  // convergence test (no preconditioner runs for nothing after the last one)
  auto body = [&](hipStream_t st, int m, int par) {
    for (int k = 0; k < m; ++k) {
      const int i = par + k;   // parity of the global iteration index
      z32 = correction(0, st);
      iteration(pbuf[i & 1], pbuf[(i + 1) & 1], st);
    }
  };
  // iteration 0 ran eagerly: the chunks start at an odd index
  int rc = run_chunks(P, 1, jacobi, maxit, 1, body, s);
  if (rc) return rc;
  if (P.launch_refused) return 2;   // tk_last_error() names the sweep
  if (hipGetLastError() != hipSuccess) {
    set_error("kernel launch failed in mixed pcg");
    return 2;
  }
  return 0;
}

// ------------------------------------------------------------------ hierarchy
int setup_hierarchy(Plan& P, const double* rho, double penal, int prec, hipStream_t s) {
  Level& L0 = P.lv[0];
  // FP64 scales in either precision (the FP64 outer operator and the coarsest
  // factorisation read them); the FP32 copy from the same pass
  if (prec != 0) launch_stiffness_scale2(P, rho, penal, L0.f64.s, L0.f32.s, s);
  else launch_stiffness_scale<double>(P, rho, penal, L0.f64.s, s);
  if (prec != 0) {
    launch_coarsen_scale2(P, s);
  } else {
    for (int l = 0; l + 1 < P.nlev; ++l)
      launch_coarsen_scale<double>(P.lv[l], P.lv[l + 1], P.lv[l].f64.s, P.lv[l + 1].f64.s, s);
  }
  // the Galerkin level's coefficients from its children's scales (the
  // rediscretised scales of that level still feed the levels below it)
  if (P.gal_level > 0) {
    if (prec == 0) launch_gal_coeffs<double>(P, s);
    else launch_gal_coeffs<float>(P, s);
  }
  // The coarsest factorisation (one workgroup, ~0.4 ms) runs on the plan's
  // set-up stream beside the inverse diagonals, the right-hand side, the
  // initial residual and the first V-cycle's descent; the first coarsest solve
  // waits for it (join_coarse).  The set-up stream starts after everything
  // queued on s so far, so it never overwrites a factor still in use.
  TK_HIP(hipEventRecord(P.cev_in, s));
  TK_HIP(hipStreamWaitEvent(P.cstream, P.cev_in, 0));
  launch_coarse_setup(P, P.cstream);
  TK_HIP(hipEventRecord(P.cev_out, P.cstream));
  P.coarse_pending = true;
  // inverse diagonals for the V-cycle (and Jacobi) of this precision only
  if (prec == 0) launch_diag_inv_levels<double>(P, s);
  else launch_diag_inv_levels<float>(P, s);
  if (P.gal_level > 0) {
    if (prec == 0) launch_gal_diag<double>(P, s);
    else launch_gal_diag<float>(P, s);
  }
  TK_HIP(hipGetLastError());
  P.last_penal = penal;
  P.last_prec = prec;
  P.last_rho = rho;
  P.setup_mask_gen = P.mask_gen;
  return 0;
}

// The mixed solver's direction ring (ring_m FP32 vectors and their step
// lengths) is allocated by the first mixed solve with the element operator,
// before any mixed chunk is captured, so FP64-only plans never hold it.
static int ensure_ring(Plan& P) {
  if (P.ring_tried || !P.use_el) return 0;
  P.ring_tried = true;
  const long long n = P.g.ndof;
  if (hipMalloc((void**)&P.pring, sizeof(float) * (size_t)P.ring_m * (size_t)n) != hipSuccess ||
      hipMalloc((void**)&P.alpha_hist, sizeof(double) * kRingMax) != hipSuccess) {
    (void)hipGetLastError();
    if (P.pring) (void)hipFree(P.pring);
    P.pring = nullptr;
    set_error("could not allocate the mixed solver's direction ring (" +
              std::to_string(P.ring_m) + " FP32 vectors of " + std::to_string(n) +
              " entries); lower direction_slots");
    return 2;
  }
  return 0;
}

// After an expired barrier wait of the persistent coarse cycle (its workgroups
// did not all run on one XCD at once): the plan runs those levels a launch per
// step from now on.  The captured chunks hold the persistent launches, so they
// are dropped; the error flag and the barrier words are cleared.
static int drop_pcycle(Plan& P, hipStream_t s) {
  P.pc_nlev = 0;
  P.pc_first = -1;
  for (auto& a : P.gexec)
    for (auto& b2 : a)
      for (auto& c : b2)
        for (auto& ex : c)
          if (ex) {
            (void)hipGraphExecDestroy(ex);
            ex = nullptr;
          }
  const int zero = 0;
  TK_HIP(hipMemcpyAsync(&P.sc->pc_err, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  TK_HIP(hipMemsetAsync(P.pc_sync, 0, sizeof(unsigned) * kPcSyncWords, s));
  TK_HIP(hipStreamSynchronize(s));
  return 0;
}

// ------------------------------------------------------------------ driver
// flags: bit0 = U holds the initial guess, bit1 = Jacobi instead of multigrid,
//        bit2 = skip hierarchy setup (reuse the last one for the same rho).
int solve(Plan& P, int prec, const double* rho, double penal, const double* F, double* U,
          double rtol, int maxit, int flags, int* iters_out, double* relres_out, hipStream_t s) {
  const int n = P.g.ndof;
  const int use_x0 = flags & 1, jacobi = (flags >> 1) & 1, reuse = (flags >> 2) & 1;
  if (prec != 0) {
    if (int rc = ensure_ring(P)) return rc;
  }
  if (!reuse) {
    int rc = setup_hierarchy(P, rho, penal, prec, s);
    if (rc) return rc;
  }
  P.cnt.solves++;
  P.launch_refused = false;
  double* b = P.lv[0].f64.b;  // masked right-hand side
  // the iterate lives in the plan (x64) so that captured chunks stay valid
  // for every caller buffer; U is copied in (warm start, masked in the same
  // pass as the right-hand side) and out
  double* x = P.x64;
  hipLaunchKernelGGL(k_mask_rhs, dim3(nblocks(P.g.nnode)), dim3(kBlock), 0, s, P.g.nnode,
                     P.lv[0].fixed, F, b, use_x0 ? (const double*)U : nullptr, x);
  int total_it = 0;
  double relres = 0.0;
  int rc = (prec == 0) ? pcg<double>(P, b, x, P.r64, P.p64, P.q64, !use_x0, rtol, maxit, jacobi, s)
                       : pcg_mixed(P, b, x, !use_x0, rtol, maxit, jacobi, s);
  if (rc) return rc;
  if (P.sc_host->pc_err) {
    if (int e = drop_pcycle(P, s)) return e;
    set_error("the persistent coarse cycle timed out at a barrier; the plan now launches those "
              "levels step by step: solve again");
    return 3;
  }
  const int c = (prec != 0 && P.use_el && P.pring) ? P.sc_host->iter % P.ring_m : 0;
  if (c > 0)
    hipLaunchKernelGGL(k_ring_fold_out, dim3(nblocks(n)), dim3(kBlock), 0, s, n, x, P.pring, c,
                       P.alpha_hist, U);
  else
    TK_HIP(hipMemcpyAsync(U, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (prec == 0) {
    total_it = P.sc_host->iter;
    relres = (P.sc_host->bb > 0.0) ? std::sqrt(P.sc_host->rr / P.sc_host->bb) : 0.0;
    P.cnt.cg_iters += total_it;
    if (P.sc_host->stalled && !(relres <= rtol)) {
      if (iters_out) *iters_out = total_it;
      if (relres_out) *relres_out = relres;
      set_error("CG failed to converge: breakdown at iteration " + std::to_string(total_it));
      return 3;
    }
  } else {
    total_it = P.sc_host->iter;
    relres = (P.sc_host->bb > 0.0) ? std::sqrt(P.sc_host->rr / P.sc_host->bb) : 0.0;
    P.cnt.ir_steps += total_it;
    P.cnt.cg_iters += total_it;
  }
  if (iters_out) *iters_out = total_it;
  if (relres_out) *relres_out = relres;
  if (!(relres <= rtol)) {
    char buf[160];
    snprintf(buf, sizeof(buf), "CG failed to converge: relative residual %.3e (target %.3e) after %d iterations",
             relres, rtol, total_it);
    set_error(buf);
    return 3;
  }
  TK_HIP(hipGetLastError());
  return 0;
}

// Public helper used by the API for a single operator application.
// y = A_l u with multigrid level l's operator as the V-cycle applies it (the
// scales, or on the Galerkin level its coefficients, of the last set-up).
int apply_level(Plan& P, int prec, int l, const void* u, void* y, hipStream_t s) {
  if (prec == 0)
    launch_apply<double>(P, P.lv[l], P.lv[l].f64.s, (const double*)u, (double*)y, nullptr, 0, s);
  else
    launch_apply<float>(P, P.lv[l], P.lv[l].f32.s, (const float*)u, (float*)y, nullptr, 0, s);
  TK_HIP(hipGetLastError());
  return 0;
}

// x = the V-cycle's single-workgroup part on the tail level for right-hand side
// b (zero on the level's fixed DOFs), with the smoothing coefficients the
// V-cycle uses there.
int apply_tail(Plan& P, int prec, const void* b, void* x, hipStream_t s) {
  if (P.tail_level < 1) {
    set_error("this plan's V-cycle has no single-workgroup tail");
    return 1;
  }
  join_coarse(P, s);
  const TailCoef cf = tail_coefs(P, P.tail_level);
  if (prec == 0)
    launch_tail<double>(P, (const double*)b, (double*)x, cf, s, nullptr);
  else
    launch_tail<float>(P, (const float*)b, (float*)x, cf, s, nullptr);
  TK_HIP(hipGetLastError());
  return 0;
}

// x = the persistent coarse cycle with the tail below it (the V-cycle's part
// from level pc_first down and back) for right-hand side b on level pc_first.
int apply_coarse_cycle(Plan& P, int prec, const void* b, void* x, hipStream_t s) {
  if (P.pc_nlev < 1) {
    set_error("this plan's V-cycle has no persistent coarse levels");
    return 1;
  }
  const int l = P.pc_first, lt = P.tail_level;
  const size_t n = (size_t)P.lv[l].g.ndof;
  const int zero = 0;   // the kernels return at once when the solver's stop flag is set
  TK_HIP(hipMemcpyAsync(&P.sc->done, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  join_coarse(P, s);
  const TailCoef cf = tail_coefs(P, lt);
  if (prec == 0) {
    TK_HIP(hipMemcpyAsync(P.lv[l].f64.b, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    launch_pcycle_down<double>(P, cf, s);
    launch_tail<double>(P, P.lv[lt].f64.b, P.lv[lt].f64.x, cf, s, nullptr);
    launch_pcycle_up<double>(P, cf, s);
    TK_HIP(hipMemcpyAsync(x, P.lv[l].f64.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  } else {
    TK_HIP(hipMemcpyAsync(P.lv[l].f32.b, b, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    launch_pcycle_down<float>(P, cf, s);
    launch_tail<float>(P, P.lv[lt].f32.b, P.lv[lt].f32.x, cf, s, nullptr);
    launch_pcycle_up<float>(P, cf, s);
    TK_HIP(hipMemcpyAsync(x, P.lv[l].f32.x, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  TK_HIP(hipGetLastError());
  TK_HIP(hipStreamSynchronize(s));
  TK_HIP(hipMemcpy(P.sc_host, P.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  if (P.sc_host->pc_err) {
    if (int e = drop_pcycle(P, s)) return e;
    set_error("the persistent coarse cycle timed out at a barrier; the plan now launches those "
              "levels step by step");
    return 3;
  }
  return 0;
}

int apply_fine(Plan& P, int prec, const void* u, void* y, hipStream_t s) {
  if (prec == 0)
    launch_apply<double>(P, P.lv[0], P.lv[0].f64.s, (const double*)u, (double*)y, nullptr, 0, s);
  else
    launch_apply<float>(P, P.lv[0], P.lv[0].f32.s, (const float*)u, (float*)y, nullptr, 0, s);
  TK_HIP(hipGetLastError());
  return 0;
}

// One fine-level Chebyshev smoothing step with the scales and inverse diagonal
// of the last solve: d = c1 d + c2 D^-1 (b - K x); xo = x + d.
int smooth_step_fine(Plan& P, int prec, const void* b, const void* x, void* xo, void* d,
                     double c1, double c2, hipStream_t s) {
  Level& L = P.lv[0];
  if (!P.use_el) {
    set_error("tk_smoother_step needs the element operator (mirror-symmetric KE)");
    return 1;
  }
  if (prec == 0)
    el_cheb<double>(P, L, L.f64.s, L.f64.dinv, (const double*)b, (const double*)x, (double*)xo,
                    (double*)d, 0.0, c1, c2, 0, 1, nullptr, s);
  else
    el_cheb<float>(P, L, L.f32.s, L.f32.dinv, (const float*)b, (const float*)x, (float*)xo,
                   (float*)d, 0.0, c1, c2, 0, 1, nullptr, s);
  TK_HIP(hipGetLastError());
  return 0;
}

}  // namespace tk
