/*
 * pytopo3d_hip.h -- C ABI of the MI355X (gfx950) SIMP hot-path library
 * (libpytopo3d_hip.so).
 *
 * It replaces the per-iteration arithmetic of PyTopo3D's
 * pytopo3d.core.optimizer.top3d (pytopo3d/core/optimizer.py:578-696).  The
 * reference has no FFI of its own; every entry point below names the Python
 * function or call site it stands in for; pytopo3d_amd/_native.py is the
 * ctypes binding.
 *
 * Conventions
 *   - Array arguments are DEVICE pointers (contiguous, caller-owned), except
 *     where a parameter is documented as host memory.
 *   - Element arrays: nele = nelx*nely*nelz doubles, element e = iy + ix*nely
 *     + iz*nelx*nely (Fortran order of the (nely, nelx, nelz) design array).
 *   - DOF arrays: ndof = 3*(nelx+1)*(nely+1)*(nelz+1), DOF 3*n + c of node
 *     n = iy + ix*(nely+1) + iz*(nelx+1)*(nely+1), c = x, y, z.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Work is enqueued
 *     on it; calls that return a host scalar synchronise it.
 *   - Return value: 0 ok, 1 invalid argument (the reference raises
 *     ValueError), 2 HIP/runtime error, 3 linear solve not converged (the
 *     reference raises RuntimeError, pytopo3d/utils/solver.py:57-58).
 *     tk_last_error() gives the message of the last failure on this thread.
 *   - A plan is not thread-safe; distinct plans are independent.  Calls on a
 *     plan allocate no device memory (the workspaces are made by
 *     tk_plan_create), except the first mixed-precision tk_solve on a plan,
 *     which allocates the solver's direction ring (direction_slots FP32
 *     vectors, about 2 GiB by default), and the first tk_solve of each
 *     precision, which creates its solver stream and graphs.
 */
#ifndef PYTOPO3D_HIP_H
#define PYTOPO3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tk_plan tk_plan;

/* Precision of the linear solve / operator. */
enum {
  TK_FP64 = 0,       /* FP64 operator, FP64 PCG */
  TK_MIXED_FP32 = 1  /* FP32 operator and multigrid inside FP64 iterative refinement */
};

/* tk_solve flags */
enum {
  TK_SOLVE_X0 = 1,        /* U holds the initial guess (warm start) */
  TK_SOLVE_JACOBI = 2,    /* Jacobi instead of the multigrid preconditioner */
  TK_SOLVE_REUSE = 4      /* reuse the hierarchy built by the last solve or tk_prepare: the
                             same rho buffer (unmodified: the contents are not compared),
                             penal and precision, and no tk_plan_set_masks since; anything
                             else is rejected with status 1 */
};

typedef struct tk_options {
  int smooth_degree;      /* Chebyshev smoothing steps per pre/post smoothing (default 2) */
  double cheb_lo;         /* smoothing interval, fractions of the lambda_max bound (0.06) */
  double cheb_hi;         /* (1.0) */
  int max_levels;         /* multigrid levels (12) */
  int coarse_max_dof;     /* coarsest grid DOF limit, <= 192 (192) */
  int check_every;        /* iterations between host convergence checks (8) */
  int operator_kind;      /* 0: element kernel in the Walsh-Hadamard sector form when KE is
                             mirror-symmetric (default), 1: dense node-gather kernel,
                             2: as 0, with the FP32 sweeps (the mixed solve's multigrid,
                             tk_apply_K and tk_smoother_step in TK_MIXED_FP32) contracting
                             each element as a BF16 [elements x 32] x [32 x 32] MFMA
                             product with FP32 accumulation (FP64 sweeps unchanged) */
  int no_graphs;          /* 0: each chunk of check_every Krylov iterations is captured once
                             per plan as a HIP graph and replayed (default); 1: kernels are
                             launched one by one */
  int filter_kind;        /* density filter kernel: 0 LDS tiles when the stencil radius is
                             <= 3 (default), 1 stencil gather (the only form for larger
                             radii) */
  int coarse_degree;      /* Chebyshev smoothing steps per pre/post smoothing on the third
                             and coarser multigrid levels of at most 32768 elements (default
                             12; those the persistent coarse cycle and the single-workgroup
                             tail run); smooth_degree applies to the two finest, mid_degree
                             to the coarse levels of more elements */
  double coarse_cheb_lo;  /* low end of the smoothing interval on those levels (0.0005):
                             they are latency-bound, so a longer, wider smoother costs
                             little there and brings the cycle closer to a coarse solve */
  int direction_slots;    /* mixed solves with the element kernel: FP32 search directions
                             kept (2..64) so that the iterate is updated once per that many
                             iterations and at the end rather than in every update pass;
                             0 (default): as many as fit in 2 GiB, at most 64.  The result
                             does not depend on it (same sums in the same order) */
  int galerkin;           /* coarse operator of the third multigrid level (coarse elements of
                             4x4x4 fine elements): 1 the Galerkin product P^T A P of the
                             rediscretised second level (exact per coarse element: its 8
                             children's scales enter separately, not as their mean); 0
                             (default) or 2 rediscretised like the other levels.  Used when
                             that level is neither the coarsest nor part of the single-
                             workgroup tail.  It saves CG iterations on designs with sharp
                             solid/void contrast inside 4x4x4 blocks, costs ~5 % per
                             iteration, and left the iteration counts of smooth SIMP designs
                             (256x128x128, 320x160x160) unchanged; the solve's tolerance is
                             the same either way */
  int persistent_coarse;  /* 1 (default): the coarse levels between the third level and the
                             single-workgroup tail that split into 32 bricks of <= 1024
                             elements (32x16x16 and 16x8x8 at 256x128x128) run their whole
                             V-cycle part in one launch per direction on the 32 compute units
                             of one XCD, which meet at barriers between smoothing steps,
                             instead of a ~5 us launch per step; -1: a launch per step; 0 keeps
                             the default.  The arithmetic is the same up to rounding order */
  int mid_degree;         /* Chebyshev smoothing steps on the third and coarser levels of
                             more than 32768 elements (default 8; 64x32x32 at 256x128x128),
                             each a launch per step */
  double mid_cheb_lo;     /* low end of the smoothing interval on those levels (0.002) */
} tk_options;

/* Fill *opt with the defaults. */
void tk_options_default(tk_options* opt);

/* Create a plan for a nelx x nely x nelz grid.
 * KE: HOST pointer, 576 doubles, row-major 24x24 element stiffness in the
 *     corner order of pytopo3d/utils/assembly.py:17-26 (lk_H8,
 *     pytopo3d/utils/stiffness.py:12-130).
 * rmin: filter radius; builds the stencil form of H and Hs
 *     (replaces build_filter, pytopo3d/utils/filter.py:122-228).
 * opt: NULL for defaults. */
int tk_plan_create(tk_plan** out, int nelx, int nely, int nelz, const double* KE, double rmin,
                   const tk_options* opt);
int tk_plan_destroy(tk_plan* plan);

/* Masks (device pointers, copied into the plan):
 * elem_flags: nele bytes, bit0 = protected solid, bit1 = protected void
 *     (the masks of pytopo3d/core/optimizer.py:343-405, void wins);
 * node_fixed: nnode bytes, bit c = DOF c of the node is fixed
 *     (build_supports, pytopo3d/utils/assembly.py:224-348).
 * Returns the free-element count in *nfree (host). */
int tk_plan_set_masks(tk_plan* plan, const uint8_t* elem_flags, const uint8_t* node_fixed,
                      long long* nfree, void* stream);

/* Boundary conditions on the device from the reference's inputs, each of which
 * may be NULL (the reference's None):
 *   force_field          (nely, nelx, nelz, 3) doubles, C order (the reference's
 *                        array as is); NULL: the default line load
 *   support_mask         (nely, nelx, nelz) bytes, C order; NULL or all zero:
 *                        the x = 0 face is fixed
 *   obstacle_mask,
 *   protected_zone_mask  (nely, nelx, nelz) bytes, C order
 * Outputs (device, each may be NULL, at least one given):
 *   F          ndof doubles: build_force_vector (pytopo3d/utils/assembly.py:78-184),
 *              bitwise (each node's element contributions added in the
 *              reference's loop order)
 *   node_fixed nnode bytes as tk_plan_set_masks takes them: build_supports
 *              (pytopo3d/utils/assembly.py:224-348)
 *   elem_flags nele bytes in element order as tk_plan_set_masks takes them: the
 *              protected-solid / void masks of pytopo3d/core/optimizer.py:343-405
 *              (load and support elements or the default strips are solid,
 *              void wins). */
int tk_build_boundary_conditions(tk_plan* plan, const double* force_field,
                                 const uint8_t* support_mask, const uint8_t* obstacle_mask,
                                 const uint8_t* protected_zone_mask, double* F,
                                 uint8_t* node_fixed, uint8_t* elem_flags, void* stream);

/* Plan geometry: writes nlev and up to cap*3 element dims (nelx,nely,nelz per level). */
int tk_plan_levels(const tk_plan* plan, int* nlev, int* dims, int cap);
/* Stencil table (HOST outputs): dy,dx,dz per entry and the weight; returns count in *n. */
int tk_plan_stencil(const tk_plan* plan, int* offsets, double* weights, int cap, int* n);
/* Density-filter kernel the plan uses (HOST outputs, each may be NULL):
 * tiled = 1 for the LDS-tiled stencil (stencil reach <= 3 elements per axis,
 * filter_kind 0), 0 for the stencil gather; halo = the stencil's largest
 * per-axis offset, which is the tile halo of the tiled kernel; nsten = stencil
 * entries.  Not in the reference (its filter is a sparse matrix,
 * pytopo3d/utils/filter.py:122-228); it lets tests name the kernel variant. */
int tk_plan_filter_info(const tk_plan* plan, int* tiled, int* halo, int* nsten);
/* Galerkin level of the plan's multigrid hierarchy (HOST outputs, each may be
 * NULL): level = the level whose operator is the Galerkin product of the level
 * above (-1: none, every level rediscretised), bound = that level's bound of
 * lambda_max(D^-1 A) used by its Chebyshev smoother (KE's bound when none).
 * No counterpart in the reference (its solve is a library call,
 * pytopo3d/utils/solver.py:18-73); it lets tests name the hierarchy. */
int tk_plan_galerkin_info(const tk_plan* plan, int* level, double* bound);
/* Copy the filter row sums Hs (device output, nele doubles). */
int tk_filter_row_sums(tk_plan* plan, double* Hs, void* stream);

/* rho_f = (H x)/Hs, rho = Heaviside projection clamped to [0,1] with protected
 * solid/void forced to 1/0, drho = d rho / d rho_f (0 on protected elements).
 * Replaces build_physical_density + heaviside_projection
 * (pytopo3d/utils/filter.py:26-100).  Any output may be NULL. */
int tk_filter_project(tk_plan* plan, const double* x, double beta, double eta, double* rho_f,
                      double* rho, double* drho, void* stream);

/* dc_dx = H^T((g_c * drho)/Hs), dv_dx = H^T((g_v * drho)/Hs).
 * g_v == NULL uses the free-element indicator (dv/drho of optimizer.py:613).
 * dv_dx may be NULL (then g_v is ignored).
 * Replaces apply_density_filter_chain_rule (pytopo3d/utils/filter.py:103-119)
 * at pytopo3d/core/optimizer.py:614-625. */
int tk_filter_transpose2(tk_plan* plan, const double* g_c, const double* g_v,
                         const double* drho, double* dc_dx, double* dv_dx, void* stream);

/* y = K(rho) u on the free DOFs (identity rows on fixed DOFs), K = sum_e
 * (1e-9 + (1-1e-9) rho_e^penal) KE.  prec TK_FP64: u, y double*; TK_MIXED_FP32:
 * float*.  rho == NULL reuses the element scales of the last tk_solve on this
 * plan (a single kernel launch; used for timing).  Replaces the assembly + Kff
 * of pytopo3d/core/optimizer.py:584-591. */
int tk_apply_K(tk_plan* plan, int prec, const double* rho, double penal, const void* u, void* y,
               void* stream);

/* y = A_l u with multigrid level l's operator as the V-cycle applies it
 * (identity rows on the level's fixed DOFs), built by the last tk_solve or
 * tk_prepare in precision prec; u, y: 3 (nx_l + 1)(ny_l + 1)(nz_l + 1) values
 * (dimensions from tk_plan_levels) of that precision.  No counterpart in the
 * reference (it has no multigrid); it lets tests check each level's operator. */
int tk_level_apply(tk_plan* plan, int prec, int level, const void* u, void* y, void* stream);
/* The V-cycle's grid transfers between level `level` and the next coarser one
 * (precision prec, values of that precision, node counts from tk_plan_levels;
 * the plan's masks must be set): direction 0, out (coarse) = R in, the
 * trilinear restriction R = P^T of a fine field that is zero on the fine fixed
 * DOFs, zero on the coarse fixed DOFs; direction 1, out (fine) += P in, the
 * trilinear prolongation, added on the fine free DOFs only.  No counterpart in
 * the reference (it has no multigrid); it lets tests check the transfers. */
int tk_transfer_apply(tk_plan* plan, int prec, int level, int direction, const void* in,
                      void* out, void* stream);
/* x = the part of the V-cycle that one workgroup runs on the tail level (the
 * second-coarsest level): pre-smoothing from x = 0, residual, restriction,
 * coarsest solve with the stored inverse, prolongation and post-smoothing,
 * with the hierarchy of the last tk_solve or tk_prepare in precision prec.
 * b (zero on the level's fixed DOFs) and x: 3 (nx_l + 1)(ny_l + 1)(nz_l + 1)
 * values of that precision; fails when the plan has no tail level.  No
 * counterpart in the reference (it has no multigrid); it lets tests check the
 * tail against the level operators of tk_level_apply. */
int tk_tail_apply(tk_plan* plan, int prec, const void* b, void* x, void* stream);
/* The tail level of the plan's V-cycle (-1: none; then the coarse levels run
 * level by level down to the coarsest solve). */
int tk_plan_tail_level(const tk_plan* plan, int* level);
/* The levels of the persistent coarse cycle (tk_options.persistent_coarse):
 * *first (-1: none) and *count, the levels first .. first + count - 1 just
 * above the tail level. */
int tk_plan_persistent_levels(const tk_plan* plan, int* first, int* count);
/* How long a workgroup of the persistent coarse cycle waits at a barrier for
 * the others (default 0.05 s; [0, 10]).  When a wait expires the solve returns
 * 3 and the plan runs those levels a launch per step from then on.  No
 * counterpart in the reference; it lets tests force that fallback. */
int tk_plan_set_barrier_timeout(tk_plan* plan, double seconds);
/* x = the V-cycle's part from the first persistent level down and back (the
 * persistent launches and the single-workgroup tail between them) for
 * right-hand side b (zero on that level's fixed DOFs), with the hierarchy of
 * the last tk_solve or tk_prepare in precision prec; b, x: that level's DOF
 * count of values of the precision.  Fails when the plan has no persistent
 * levels.  No counterpart in the reference (it has no multigrid); it lets
 * tests check the persistent cycle against independently assembled level
 * operators. */
int tk_coarse_cycle_apply(tk_plan* plan, int prec, const void* b, void* x, void* stream);
/* One fine-grid Chebyshev smoothing step of the multigrid preconditioner,
 * with the element scales and inverse diagonal of the last tk_solve in the
 * same precision: d = c1 d + c2 D^-1 (b - K x); xo = x + d (identity rows on
 * fixed DOFs).  b, x, xo, d: ndof doubles (TK_FP64) or floats (TK_MIXED_FP32).
 * The smoother has no counterpart in the reference (its direct solve needs
 * none); it is exported for timing and for callers building their own cycles. */
int tk_smoother_step(tk_plan* plan, int prec, const void* b, const void* x, void* xo, void* d,
                     double c1, double c2, void* stream);
/* diag(K(rho)) (1 on fixed DOFs), ndof doubles. */
int tk_diag_K(tk_plan* plan, const double* rho, double penal, double* diag, void* stream);

/* Build the solver state of a design without solving: element scales of every
 * multigrid level, inverse diagonals and the coarsest-grid inverse, in the
 * given precision (TK_MIXED_FP32 also forms the FP64 element scales, which its
 * FP64 outer iteration reads, but no FP64 inverse diagonals).  A following
 * tk_solve with TK_SOLVE_REUSE (same rho, penal and precision), tk_apply_K with
 * rho == NULL and tk_smoother_step (same precision) use it.  This is the set-up half of the solve at
 * pytopo3d/core/optimizer.py:584-594; it lets callers with several load cases
 * per design build it once. */
int tk_prepare(tk_plan* plan, int prec, const double* rho, double penal, void* stream);

/* Solve K(rho) U = F on the free DOFs, U = 0 on fixed DOFs.
 * Replaces get_solver + the solve at pytopo3d/core/optimizer.py:591-594
 * (pytopo3d/utils/solver.py:18-73).  rtol: relative residual target; iters_out
 * and relres_out (HOST, may be NULL) receive the iteration count and the final
 * relative residual.  Returns 3 when rtol is not reached. */
int tk_solve(tk_plan* plan, int prec, const double* rho, double penal, const double* F,
             double* U, double rtol, int maxit, int flags, int* iters_out, double* relres_out,
             void* stream);

/* Element energies and sensitivities from a displacement field:
 * ce = u_e.KE.u_e (element_compliance, pytopo3d/core/compliance.py:18-75),
 * c = sum_e s_e ce (optimizer.py:601-605), dc_drho = -p(1-1e-9)rho^(p-1) ce
 * (optimizer.py:607-612).  If dc_dx/dv_dx are given (drho required) the chain
 * rule of optimizer.py:614-625 is applied too.  Outputs may be NULL.
 * c_out (HOST, may be NULL): if given the stream is synchronised. */
int tk_energy(tk_plan* plan, const double* U, const double* rho, const double* drho, double penal,
              double* ce, double* dc_drho, double* dc_dx, double* dv_dx, double* c_out,
              void* stream);

/* Allocation-free variant of tk_energy without a plan: ce, dc_drho (device
 * outputs, each may be NULL) and c = sum_e s_e ce of a displacement field U on
 * a nelx x nely x nelz grid.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  KE: HOST pointer, 576
 * doubles, with the element's mirror symmetry (every axis-aligned lk_H8 KE).
 * c_dev (device, 1 double) and/or c_host (host; synchronises the stream, so
 * not under capture) receive c; either may be NULL.
 * Limit: nelx*nely*nelz <= 16,777,216 (the library's static scratch: 64 Ki
 * partial sums plus 1 KiB of scalars, 513 KiB held on each device it runs on
 * until the process ends); larger grids return 1.  It calls no allocating
 * routine.  Every call on a device shares that scratch, which keeps the last
 * call's contents: calls that overlap on one device (several streams or host
 * threads) race and must be serialised by the caller. */
int tk_energy_noalloc(int nelx, int nely, int nelz, const double* KE, const double* U,
                      const double* rho, double penal, double* ce, double* dc_drho,
                      double* c_dev, double* c_host, void* stream);

/* Projected-volume OC update (optimality_criteria_update_projected,
 * pytopo3d/utils/oc_update.py:25-124), the bisection entirely on the device.
 * x, dc_dx, dv_dx: nele doubles; x_new: output.  target: free-volume target
 * (a sum over free elements).  steps_out (HOST, may be NULL): the bisection
 * steps, as the reference counts them.  Synchronises the stream: the host
 * reads the device bisection state between chunks of volume evaluations, which
 * stop when the bisection has (midpoints that an earlier evaluation decides
 * through the monotone volume are not measured; the result is the reference's
 * bisection result). */
int tk_oc_update(tk_plan* plan, const double* x, const double* dc_dx, const double* dv_dx,
                 double move, double target, double beta, double eta, int max_steps, double tol,
                 double* x_new, int* steps_out, void* stream);

/* Statistics over free elements (pytopo3d/core/optimizer.py:642-655):
 * change = max|x_new - x_old| (x_old/x_new may be NULL: change = -inf),
 * vol_sum = sum rho, gray = #(0.05 < rho < 0.95).  Left on the device; read
 * with tk_read_results. */
int tk_design_stats(tk_plan* plan, const double* x_old, const double* x_new, const double* rho,
                    void* stream);

/* Mean free physical density of the all-void and all-solid free designs at
 * beta (achievable_free_volume_bounds, pytopo3d/core/optimizer.py:501-531).
 * HOST outputs; synchronises. */
int tk_volume_bounds(tk_plan* plan, double beta, double eta, double* vmin, double* vmax,
                     void* stream);

/* Synchronise and copy the device results (HOST out[0..n)):
 * [0] c, [1] change, [2] vol_sum, [3] gray count, [4] OC steps, [5] OC lambda,
 * [6] OC volume (of the last evaluation), [7] OC ratio scale, [8] nfree,
 * [9] OC volume evaluations. */
int tk_read_results(tk_plan* plan, double* out, int n, void* stream);

/* Work counters (HOST out[0..n)): [0] FP64 fine operator applications,
 * [1] FP32 fine applications, [2] Krylov iterations, [3] refinement steps,
 * [4] solves, [5] OC bisection steps, [6] filter passes. */
int tk_plan_counters(const tk_plan* plan, long long* out, int n);

/* Message of the last failure on the calling thread ("" if none). */
const char* tk_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PYTOPO3D_HIP_H */
