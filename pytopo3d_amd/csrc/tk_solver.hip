// Linear solve K U = F on the free DOFs: preconditioned conjugate gradients
// with a geometric-multigrid V-cycle (or Jacobi), in FP64, or as FP64
// iterative refinement around an FP32 inner solve.
//
// Replaces the solver factory of pytopo3d/utils/solver.py:18-73 (PyPardiso /
// SuperLU direct solve, CuPy CG+Jacobi) at its call site
// pytopo3d/core/optimizer.py:591-594.  All step lengths, dot products and the
// refinement residual are FP64; the host only reads the device scalars every
// Options::check_every iterations.
#include "tk_plan.hpp"

#include <cmath>
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

// sum (a.a) and (b.b) -> sc->rr, sc->bb; arms the iteration.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_norms(int n, const T* __restrict__ b,
                                                      const T* __restrict__ r, double tol2,
                                                      int maxit, double* partials,
                                                      unsigned* counter, Scalars* sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double bi = b[i], ri = r[i];
    v[0] = bi * bi;
    v[1] = ri * ri;
  }
  grid_sum<2>(v, partials, counter, [&](double (&t)[2]) {
    sc->bb = t[0];
    sc->rr = t[1];
    sc->rr0 = t[1];
    sc->tol2 = tol2;
    sc->maxit = maxit;
    sc->iter = 0;
    sc->stalled = 0;
    sc->done = (t[1] <= tol2 * t[0]) ? 1 : 0;
    if (!(t[1] == t[1])) { sc->done = 1; sc->stalled = 1; }
  });
}

// x += alpha p; r -= alpha q; rr -> convergence test.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int n, T* __restrict__ x, T* __restrict__ r,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const int* __restrict__ done,
                                                       double* partials, unsigned* counter,
                                                       Scalars* sc) {
  if (*done) return;
  const T alpha = (T)sc->alpha;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] = (double)ri * (double)ri;
  }
  grid_sum<1>(v, partials, counter, [&](double (&t)[1]) {
    sc->rr = t[0];
    sc->iter += 1;
    if (t[0] <= sc->tol2 * sc->bb) sc->done = 1;
    if (sc->iter >= sc->maxit) sc->done = 1;
    if (!(t[0] == t[0])) { sc->done = 1; sc->stalled = 1; }
  });
}

// rz = r.z; beta = rz / rz_old (first: beta = 0).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_rz(int n, const T* __restrict__ r,
                                                   const T* __restrict__ z, int first,
                                                   const int* __restrict__ done, double* partials,
                                                   unsigned* counter, Scalars* sc) {
  if (*done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) v[0] = (double)r[i] * (double)z[i];
  grid_sum<1>(v, partials, counter, [&](double (&t)[1]) {
    if (first) {
      sc->beta = 0.0;
    } else {
      sc->beta = t[0] / sc->rz;
    }
    sc->rz = t[0];
    if (!(t[0] > 0.0)) { sc->done = 1; sc->stalled = (t[0] == 0.0) ? 0 : 1; }
  });
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

// r32 = r64 / ||r64||  (norm taken from sc->rr)
__global__ __launch_bounds__(kBlock) void k_ir_scale_r(int n, const double* __restrict__ r64,
                                                        float* __restrict__ r32,
                                                        const Scalars* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double inv = 1.0 / sqrt(sc->norm0);
  r32[i] = (float)(r64[i] * inv);
}

// x64 += ||r|| d32
__global__ __launch_bounds__(kBlock) void k_ir_correct(int n, double* __restrict__ x,
                                                        const float* __restrict__ d,
                                                        const Scalars* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += sqrt(sc->norm0) * (double)d[i];
}

// ||v||^2 -> norm{slot}
__global__ __launch_bounds__(kBlock) void k_norm2(int n, const double* __restrict__ v, int slot,
                                                   double* partials, unsigned* counter,
                                                   Scalars* sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double t[1] = {0.0};
  if (i < n) t[0] = v[i] * v[i];
  grid_sum<1>(t, partials, counter, [&](double (&s)[1]) {
    if (slot == 0) sc->norm0 = s[0]; else sc->norm1 = s[0];
  });
}

// F with fixed DOFs zeroed (the reduced system's right-hand side).
__global__ __launch_bounds__(kBlock) void k_mask_rhs(int nnode, const uint8_t* __restrict__ fixed,
                                                      const double* __restrict__ F,
                                                      double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * n + c] = ((f >> c) & 1) ? 0.0 : F[3 * n + c];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double* __restrict__ a,
                                                    float* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ __launch_bounds__(kBlock) void k_zero_fixed(int nnode, const uint8_t* __restrict__ fixed,
                                                        double* __restrict__ U) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const uint8_t f = fixed[n];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if ((f >> c) & 1) U[3 * n + c] = 0.0;
}

// ------------------------------------------------------------------ helpers
template <typename T> static LevelBuf<T>& buf(Level& L);
template <> LevelBuf<double>& buf<double>(Level& L) { return L.f64; }
template <> LevelBuf<float>& buf<float>(Level& L) { return L.f32; }

static int read_scalars(Plan& P, hipStream_t s) {
  TK_HIP(hipMemcpyAsync(P.sc_host, P.sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  TK_HIP(hipStreamSynchronize(s));
  return 0;
}

// Chebyshev smoothing of level l: `deg` steps from x (x_zero: from 0).
// Returns the buffer holding the result (x or its ping-pong partner).
template <typename T>
static T* smooth(Plan& P, int l, const T* b, T* x, T* xa, int x_zero, hipStream_t s) {
  Level& L = P.lv[l];
  LevelBuf<T>& B = buf<T>(L);
  const double lmax = P.lam_bound * P.opt.cheb_hi, lmin = P.lam_bound * P.opt.cheb_lo;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  launch_cheb_first<T>(P, L, B.s, B.dinv, b, x, xa, B.d, 1.0 / theta, x_zero, s);
  std::swap(x, xa);
  for (int k = 1; k < P.opt.smooth_degree; ++k) {
    const double rho_new = 1.0 / (2.0 * sigma - rho_old);
    launch_cheb_step<T>(P, L, B.s, B.dinv, b, x, xa, B.d, rho_new * rho_old, 2.0 * rho_new / delta,
                        s);
    std::swap(x, xa);
    rho_old = rho_new;
  }
  return x;
}

// z = V-cycle(b) on the hierarchy; returns the pointer holding z (a level-0 buffer).
template <typename T>
static T* vcycle(Plan& P, const T* b0, hipStream_t s) {
  const int nl = P.nlev;
  const int* done = &P.sc->done;
  T* cur[kMaxLevels];
  T* alt[kMaxLevels];
  const T* rhs[kMaxLevels];
  rhs[0] = b0;
  for (int l = 0; l < nl - 1; ++l) {
    LevelBuf<T>& B = buf<T>(P.lv[l]);
    T* x = smooth<T>(P, l, rhs[l], B.x, B.xa, 1, s);
    cur[l] = x;
    alt[l] = (x == B.x) ? B.xa : B.x;
    launch_residual<T>(P, P.lv[l], B.s, rhs[l], x, B.r, s, done);
    LevelBuf<T>& C = buf<T>(P.lv[l + 1]);
    launch_restrict<T>(P.lv[l], P.lv[l + 1], B.r, C.b, s, done);
    rhs[l + 1] = C.b;
  }
  {
    LevelBuf<T>& C = buf<T>(P.lv[nl - 1]);
    launch_coarse_solve<T>(P, rhs[nl - 1], C.x, s);
    cur[nl - 1] = C.x;
  }
  for (int l = nl - 2; l >= 0; --l) {
    launch_prolong_add<T>(P.lv[l + 1], P.lv[l], cur[l + 1], cur[l], s, done);
    cur[l] = smooth<T>(P, l, rhs[l], cur[l], alt[l], 0, s);
  }
  return cur[0];
}

template <typename T>
static const T* precondition(Plan& P, const T* r, T* zbuf, int jacobi, hipStream_t s) {
  if (jacobi || P.nlev == 1) {
    if (P.nlev == 1 && !jacobi) {
      launch_coarse_solve<T>(P, r, zbuf, s);
      return zbuf;
    }
    const int n = P.g.ndof;
    hipLaunchKernelGGL(k_jacobi<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, buf<T>(P.lv[0]).dinv,
                       r, zbuf, &P.sc->done);
    return zbuf;
  }
  return vcycle<T>(P, r, s);
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
  hipLaunchKernelGGL(k_cg_norms<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, b, r, rtol * rtol,
                     maxit, P.partials, P.counter, P.sc);
  // z buffer: q is free until the first operator application
  const T* z = precondition<T>(P, r, q, jacobi, s);
  hipLaunchKernelGGL(k_cg_rz<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, r, z, 1, done,
                     P.partials, P.counter, P.sc);
  hipLaunchKernelGGL(k_cg_dir<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, z, p, 1, done, P.sc);
  int it = 0;
  const int chunk = P.opt.check_every > 0 ? P.opt.check_every : 8;
  while (it < maxit) {
    const int m = (maxit - it < chunk) ? (maxit - it) : chunk;
    for (int k = 0; k < m; ++k) {
      launch_apply<T>(P, L0, B0.s, p, q, nullptr, 1, s);
      hipLaunchKernelGGL(k_cg_update<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, x, r, p, q, done,
                         P.partials, P.counter, P.sc);
      // the preconditioner output must not alias q (read by nothing after the update)
      const T* zz = precondition<T>(P, r, q, jacobi, s);
      hipLaunchKernelGGL(k_cg_rz<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, r, zz, 0, done,
                         P.partials, P.counter, P.sc);
      hipLaunchKernelGGL(k_cg_dir<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, zz, p, 0, done,
                         P.sc);
    }
    it += m;
    if (read_scalars(P, s)) return 2;
    if (P.sc_host->done) break;
  }
  if (hipGetLastError() != hipSuccess) {
    set_error("kernel launch failed in pcg");
    return 2;
  }
  return 0;
}

// ------------------------------------------------------------------ hierarchy
int setup_hierarchy(Plan& P, const double* rho, double penal, int prec, hipStream_t s) {
  Level& L0 = P.lv[0];
  launch_stiffness_scale<double>(P, rho, penal, L0.f64.s, s);
  if (prec != 0) launch_stiffness_scale<float>(P, rho, penal, L0.f32.s, s);
  for (int l = 0; l + 1 < P.nlev; ++l) {
    launch_coarsen_scale<double>(P.lv[l], P.lv[l + 1], P.lv[l].f64.s, P.lv[l + 1].f64.s, s);
    if (prec != 0)
      launch_coarsen_scale<float>(P.lv[l], P.lv[l + 1], P.lv[l].f32.s, P.lv[l + 1].f32.s, s);
  }
  for (int l = 0; l < P.nlev; ++l) {
    Level& L = P.lv[l];
    launch_diag_inv<double>(P, L, L.f64.s, L.f64.dinv, (double*)nullptr, s);
    if (prec != 0) launch_diag_inv<float>(P, L, L.f32.s, L.f32.dinv, (float*)nullptr, s);
  }
  launch_coarse_setup(P, s);
  TK_HIP(hipGetLastError());
  P.last_penal = penal;
  return 0;
}

// ------------------------------------------------------------------ driver
// flags: bit0 = U holds the initial guess, bit1 = Jacobi instead of multigrid,
//        bit2 = skip hierarchy setup (reuse the last one for the same rho).
int solve(Plan& P, int prec, const double* rho, double penal, const double* F, double* U,
          double rtol, int maxit, int flags, int* iters_out, double* relres_out, hipStream_t s) {
  const int n = P.g.ndof;
  const int use_x0 = flags & 1, jacobi = (flags >> 1) & 1, reuse = (flags >> 2) & 1;
  if (!reuse) {
    int rc = setup_hierarchy(P, rho, penal, prec, s);
    if (rc) return rc;
  }
  P.cnt.solves++;
  double* b = P.lv[0].f64.b;  // masked right-hand side
  hipLaunchKernelGGL(k_mask_rhs, dim3(nblocks(P.g.nnode)), dim3(kBlock), 0, s, P.g.nnode,
                     P.lv[0].fixed, F, b);
  if (use_x0)
    hipLaunchKernelGGL(k_zero_fixed, dim3(nblocks(P.g.nnode)), dim3(kBlock), 0, s, P.g.nnode,
                       P.lv[0].fixed, U);
  int total_it = 0;
  double relres = 0.0;
  if (prec == 0) {
    int rc = pcg<double>(P, b, U, P.r64, P.p64, P.q64, !use_x0, rtol, maxit, jacobi, s);
    if (rc) return rc;
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
    // FP64 iterative refinement around an FP32 PCG solve.
    if (!use_x0) hipLaunchKernelGGL(k_fill<double>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, U, 0.0);
    hipLaunchKernelGGL(k_norm2, dim3(nblocks(n)), dim3(kBlock), 0, s, n, b, 1, P.partials,
                       P.counter, P.sc);
    int outer = 0;
    const int max_outer = 60;
    for (;;) {
      launch_residual<double>(P, P.lv[0], P.lv[0].f64.s, b, U, P.r64, s, nullptr);
      hipLaunchKernelGGL(k_norm2, dim3(nblocks(n)), dim3(kBlock), 0, s, n, P.r64, 0, P.partials,
                         P.counter, P.sc);
      if (read_scalars(P, s)) return 2;
      const double bb = P.sc_host->norm1, rr = P.sc_host->norm0;
      relres = (bb > 0.0) ? std::sqrt(rr / bb) : 0.0;
      if (!(relres == relres)) {
        set_error("CG failed to converge: non-finite residual in refinement");
        return 3;
      }
      if (relres <= rtol || rr == 0.0 || outer >= max_outer || total_it >= maxit) break;
      hipLaunchKernelGGL(k_ir_scale_r, dim3(nblocks(n)), dim3(kBlock), 0, s, n, P.r64,
                         P.lv[0].f32.b, P.sc);
      // inner tolerance: no tighter than needed for the outer target
      double in_tol = P.opt.inner_rtol;
      const double need = 0.5 * rtol / relres;
      if (need > in_tol) in_tol = need;
      if (in_tol > 0.5) in_tol = 0.5;
      int inner_max = P.opt.inner_maxit;
      if (inner_max > maxit - total_it) inner_max = maxit - total_it;
      int rc = pcg<float>(P, P.lv[0].f32.b, P.x32, P.r32, P.p32, P.q32, 1, in_tol, inner_max,
                          jacobi, s);
      if (rc) return rc;
      total_it += P.sc_host->iter;
      hipLaunchKernelGGL(k_ir_correct, dim3(nblocks(n)), dim3(kBlock), 0, s, n, U, P.x32, P.sc);
      ++outer;
      P.cnt.ir_steps++;
    }
    P.cnt.cg_iters += total_it;
  }
  if (iters_out) *iters_out = total_it;
  if (relres_out) *relres_out = relres;
  if (!(relres <= rtol)) {
    set_error("CG failed to converge: relative residual " + std::to_string(relres) + " after " +
              std::to_string(total_it) + " iterations");
    return 3;
  }
  TK_HIP(hipGetLastError());
  return 0;
}

// Public helper used by the API for a single operator application.
int apply_fine(Plan& P, int prec, const void* u, void* y, hipStream_t s) {
  if (prec == 0)
    launch_apply<double>(P, P.lv[0], P.lv[0].f64.s, (const double*)u, (double*)y, nullptr, 0, s);
  else
    launch_apply<float>(P, P.lv[0], P.lv[0].f32.s, (const float*)u, (float*)y, nullptr, 0, s);
  TK_HIP(hipGetLastError());
  return 0;
}

}  // namespace tk
