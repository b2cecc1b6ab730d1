"""Drop-in ``top3d`` and fixed-geometry evaluation on the HIP path.

Same signature, return contract, history keys, diagnostics keys, log lines and
error messages as pytopo3d.core.optimizer (pytopo3d/core/optimizer.py:54-839).
Every per-iteration operation -- filter and projection, the matrix-free solve,
energies and sensitivities, the chain rule, the OC bisection and the
statistics -- runs as HIP kernels of libpytopo3d_hip.so on one GPU; the host
only builds the problem (KE, loads, masks) and reads a few scalars per
iteration.

Keywords beyond the reference's (all optional):
  force_vector   full load vector in the reference's DOF numbering; elements
                 touching loaded nodes become protected solid
  precision      "fp64" (default) or "mixed" (FP32 multigrid-PCG inside FP64
                 iterative refinement)
  solver         "mg" (geometric multigrid PCG, default) or "jacobi"
  solver_rtol    relative residual of every linear solve (default 1e-12)
  solver_maxit   Krylov iteration cap per solve
  solver_options dict of tk_options fields (smoothing degree, levels, ...)
  timings        dict filled with per-phase wall times and work counters
  iteration_callback
                 called after every design iteration with a dict: beta, iter,
                 total, c, change, vol, gray, and the device tensors x (the
                 new design) and rho (the physical density the iteration
                 solved with), element order; valid until the next iteration
"""

from __future__ import annotations

import logging
import time
from typing import Any, Dict, MutableMapping, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from .problem import (
    element_flags,
    element_stiffness,
    protected_masks,
    support_node_flags,
)

logger = logging.getLogger("pytopo3d_amd.optimizer")

EMIN = 1.0e-9


def _device() -> torch.device:
    _native.lib()
    return torch.device("cuda", torch.cuda.current_device())


def _to_dev_elem(a: np.ndarray, dev) -> torch.Tensor:
    """(nely, nelx, nelz) array -> contiguous device vector in element order."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))).to(dev)


def _to_host_elem(t: torch.Tensor, shape) -> np.ndarray:
    return t.detach().cpu().numpy().reshape(shape, order="F")


class SIMPProblem:
    """One discretised problem on the device: plan, masks, loads, buffers.

    Element vectors are 1-D float64 tensors in element (Fortran) order; DOF
    vectors are float64 tensors of length ndof in the reference's numbering.
    """

    def __init__(self, nelx: int, nely: int, nelz: int, rmin: float, KE: np.ndarray,
                 F: Optional[np.ndarray], node_fixed: Optional[np.ndarray], solid: np.ndarray,
                 void: np.ndarray, solver_options: Optional[dict] = None,
                 reference_bcs: Optional[dict] = None):
        """``F`` and ``node_fixed`` are host arrays, or None with
        ``reference_bcs`` = the reference's inputs (force_field, support_mask,
        and for ``device_flags`` obstacle_mask and protected_zone_mask), from
        which the loads, supports and, with ``device_flags``, the element
        flags are built on the device (tk_build_boundary_conditions)."""
        self.dev = _device()
        self.nelx, self.nely, self.nelz = int(nelx), int(nely), int(nelz)
        self.shape = (self.nely, self.nelx, self.nelz)
        self.KE = np.asarray(KE, dtype=np.float64)
        self.plan = _native.Plan(nelx, nely, nelz, self.KE, rmin, solver_options)
        self.solid = np.asarray(solid, dtype=bool)
        self.void = np.asarray(void, dtype=bool)
        self.free = ~(self.solid | self.void)
        if reference_bcs is None:
            flags = torch.from_numpy(element_flags(self.solid, self.void)).to(self.dev)
            fixed = torch.from_numpy(np.ascontiguousarray(node_fixed, dtype=np.uint8)).to(self.dev)
            self.F = torch.from_numpy(np.ascontiguousarray(F, dtype=np.float64)).to(self.dev)
        else:
            self.F, fixed, flags = self._device_bcs(**reference_bcs)
        self.nfree = self.plan.set_masks(flags, fixed)
        if self.nfree != int(np.sum(self.free)):
            raise RuntimeError(f"device element flags give {self.nfree} free elements, the host "
                               f"masks {int(np.sum(self.free))}")
        self.U = torch.zeros(self.plan.ndof, dtype=torch.float64, device=self.dev)
        self.have_U = False
        self.last_iters = 0
        self.last_relres = 0.0

    def _device_bcs(self, force_field=None, support_mask=None, obstacle_mask=None,
                    protected_zone_mask=None, device_flags=False):
        def up(a, dtype):
            if a is None:
                return None
            a = np.asarray(a, dtype=bool) if dtype is np.uint8 else np.asarray(a)
            return torch.from_numpy(np.ascontiguousarray(a.astype(dtype))).to(self.dev)

        F = torch.empty(self.plan.ndof, dtype=torch.float64, device=self.dev)
        fixed = torch.empty(self.plan.nnode, dtype=torch.uint8, device=self.dev)
        flags = (torch.empty(self.plan.nele, dtype=torch.uint8, device=self.dev) if device_flags
                 else torch.from_numpy(element_flags(self.solid, self.void)).to(self.dev))
        self.plan.boundary_conditions(
            force_field=up(force_field, np.float64), support_mask=up(support_mask, np.uint8),
            obstacle_mask=up(obstacle_mask, np.uint8) if device_flags else None,
            protected_zone_mask=up(protected_zone_mask, np.uint8) if device_flags else None,
            F=F, node_fixed=fixed, elem_flags=flags if device_flags else None)
        return F, fixed, flags

    def new_elem(self) -> torch.Tensor:
        return torch.empty(self.plan.nele, dtype=torch.float64, device=self.dev)

    def physical_density(self, x: torch.Tensor, beta: float, eta: float):
        rf, rho, d = self.new_elem(), self.new_elem(), self.new_elem()
        self.plan.filter_project(x, beta, eta, rf, rho, d)
        return rf, rho, d

    def solve(self, rho: torch.Tensor, penal: float, precision="fp64", rtol=1e-12,
              maxit=20000, solver="mg", warm=True, out: Optional[torch.Tensor] = None):
        U = self.U if out is None else out
        use_x0 = warm and self.have_U and out is None
        it, rr = self.plan.solve(rho, penal, self.F, U, rtol=rtol, maxit=maxit,
                                 precision=precision, warm_start=use_x0,
                                 jacobi=(solver == "jacobi"))
        if out is None:
            self.have_U = True
        self.last_iters, self.last_relres = it, rr
        return U

    def sensitivities(self, U, rho, drho, penal, dc_dx=None, dv_dx=None, ce=None, dc_drho=None,
                      sync=True):
        """c (sync) or None: then c stays on the device for Plan.results()."""
        c = self.plan.energy(U, rho, penal, drho=drho, ce=ce, dc_drho=dc_drho, dc_dx=dc_dx,
                             dv_dx=dv_dx, sync=sync)
        return c

    def close(self):
        self.plan.close()


class DesignState:
    """Device buffers of one design trajectory (element order).

    ``rho`` and ``rho_alt`` alternate: a design step re-filters the new design
    into ``rho_alt`` and swaps them, so after a step ``rho_alt`` still holds the
    physical density the step solved with (the reference's ``rho_fea``), ready
    for a history snapshot without a device copy.
    """

    def __init__(self, problem: SIMPProblem, x0: np.ndarray):
        self.x = _to_dev_elem(x0, problem.dev)
        self.x_new = problem.new_elem()
        self.rf, self.rho, self.drho = problem.new_elem(), problem.new_elem(), problem.new_elem()
        self.rho_alt = problem.new_elem()
        self.dc_dx, self.dv_dx = problem.new_elem(), problem.new_elem()
        self.snapshots: Optional["HistorySnapshots"] = None


class HistorySnapshots:
    """Asynchronous device-to-host copies of density snapshots.

    Each snapshot goes to its own pinned host buffer on a side stream that
    starts after the work queued so far on the caller's stream; the caller's
    stream waits for the copy only before it overwrites the source buffer
    (``fence``), so the copy overlaps the next linear solve.  ``arrays``
    synchronises and returns the snapshots as ``(nely, nelx, nelz)`` float64
    arrays (the history format of pytopo3d/core/optimizer.py:661-669).
    """

    def __init__(self, shape):
        self.shape = shape
        self.side = torch.cuda.Stream()
        self.done = None
        self.host = []

    def take(self, src: torch.Tensor) -> None:
        ready = torch.cuda.Event()
        ready.record()
        dst = torch.empty(src.numel(), dtype=src.dtype, pin_memory=True)
        with torch.cuda.stream(self.side):
            self.side.wait_event(ready)
            dst.copy_(src, non_blocking=True)
            self.done = torch.cuda.Event()
            self.done.record()
        self.host.append(dst)

    def fence(self) -> None:
        if self.done is not None:
            torch.cuda.current_stream().wait_event(self.done)
            self.done = None

    def arrays(self):
        self.side.synchronize()
        return [t.numpy().reshape(self.shape, order="F") for t in self.host]


def design_step(problem: SIMPProblem, st: DesignState, penal: float, beta: float, eta: float,
                move: float, target: float, solve_kw: dict) -> dict:
    """One design iteration of pytopo3d/core/optimizer.py:578-696 on the device:
    solve at the current rho, energies and the chain rule, the OC bisection,
    re-filtering of the new design and the free-element statistics.  The
    state's rho / drho must belong to st.x (as after filter_project)."""
    plan = problem.plan
    t0 = time.perf_counter()
    U = problem.solve(st.rho, penal, **solve_kw)
    t1 = time.perf_counter()
    # c is read with the step's other results (one wait for the device)
    problem.sensitivities(U, st.rho, st.drho, penal, dc_dx=st.dc_dx, dv_dx=st.dv_dx, sync=False)
    t2 = time.perf_counter()
    plan.oc_update(st.x, st.dc_dx, st.dv_dx, move, target, beta, eta, st.x_new)
    if st.snapshots is not None:
        st.snapshots.fence()   # rho_alt may still be copied to the host
    plan.filter_project(st.x_new, beta, eta, st.rf, st.rho_alt, st.drho)
    plan.design_stats(st.rho_alt, st.x, st.x_new)
    res = plan.results()
    t3 = time.perf_counter()
    st.x, st.x_new = st.x_new, st.x
    st.rho, st.rho_alt = st.rho_alt, st.rho
    return {"c": float(res["c"]), "change": float(res["change"]),
            "vol": res["vol_sum"] / problem.nfree,
            "gray": res["gray_cnt"] / problem.nfree, "oc_steps": int(res["oc_steps"]),
            "oc_evals": int(res["oc_evals"]),
            "solve_iters": problem.last_iters, "relres": problem.last_relres,
            "t_solve": t1 - t0, "t_sens": t2 - t1, "t_oc": t3 - t2}


def _make_problem(nelx, nely, nelz, rmin, material_params, elem_size, force_field, support_mask,
                  solid, void, force_vector, solver_options, obstacle_mask=None,
                  protected_zone_mask=None, device_flags=False) -> SIMPProblem:
    """The device problem.  Loads and supports from the reference's inputs are
    built on the device (and, with ``device_flags``, the protected-element
    flags of top3d, which ``solid`` / ``void`` mirror on the host for the
    bookkeeping); an explicit ``force_vector`` is uploaded as is."""
    KE = element_stiffness(material_params, elem_size)
    shape = (nely, nelx, nelz)
    if force_field is not None and np.shape(force_field) != shape + (3,):
        raise ValueError(f"force_field has shape {np.shape(force_field)}, but expected {shape + (3,)}")
    if support_mask is not None and np.shape(support_mask) != shape:
        raise ValueError(f"support_mask has shape {np.shape(support_mask)}, but expected {shape}")
    if force_vector is None:
        bcs = dict(force_field=force_field, support_mask=support_mask,
                   obstacle_mask=obstacle_mask, protected_zone_mask=protected_zone_mask,
                   device_flags=device_flags)
        return SIMPProblem(nelx, nely, nelz, rmin, KE, None, None, solid, void, solver_options,
                           reference_bcs=bcs)
    F = np.asarray(force_vector, dtype=float).ravel()
    if F.size != 3 * (nelx + 1) * (nely + 1) * (nelz + 1):
        raise ValueError(f"force_vector has {F.size} entries, expected {3 * (nelx + 1) * (nely + 1) * (nelz + 1)}")
    fixed = support_node_flags(nelx, nely, nelz, support_mask)
    return SIMPProblem(nelx, nely, nelz, rmin, KE, F, fixed, solid, void, solver_options)


def _fixed_geometry(problem: SIMPProblem, xPhys: np.ndarray, penal: float, precision: str,
                    rtol: float, maxit: int, solver: str) -> Tuple[float, torch.Tensor]:
    x_eval = np.array(xPhys, dtype=float, copy=True)
    x_eval[problem.solid] = 1.0
    x_eval[problem.void] = 0.0
    rho = _to_dev_elem(x_eval, problem.dev)
    U = problem.solve(rho, penal, precision=precision, rtol=rtol, maxit=maxit, solver=solver)
    c = problem.sensitivities(U, rho, None, penal)
    return c, U


def evaluate_fixed_geometry_metrics(xPhys: np.ndarray, penal: float,
                                    material_params: Optional[Tuple[float, ...]] = None,
                                    elem_size: float = 0.01,
                                    force_field: Optional[np.ndarray] = None,
                                    support_mask: Optional[np.ndarray] = None,
                                    obstacle_mask: Optional[np.ndarray] = None,
                                    protected_zone_mask: Optional[np.ndarray] = None,
                                    use_gpu: bool = False, *, force_vector=None,
                                    precision: str = "fp64", solver: str = "mg",
                                    solver_rtol: float = 1e-12, solver_maxit: int = 20000,
                                    solver_options: Optional[dict] = None,
                                    rmin: float = 1.0) -> Dict[str, Optional[float]]:
    """One FE solve on a fixed physical field (pytopo3d/core/optimizer.py:54-252).

    Returns the same keys as the reference: compliance, average load-patch
    displacements, directional equivalent stiffnesses and load totals.
    """
    nely, nelx, nelz = xPhys.shape
    shape = (nely, nelx, nelz)
    void = np.zeros(shape, bool) if obstacle_mask is None else np.asarray(obstacle_mask, bool)
    solid = (np.zeros(shape, bool) if protected_zone_mask is None
             else np.asarray(protected_zone_mask, bool))
    problem = _make_problem(nelx, nely, nelz, rmin, material_params, elem_size, force_field,
                            support_mask, solid & ~void, void, force_vector, solver_options)
    try:
        c, U = _fixed_geometry(problem, xPhys, penal, precision, solver_rtol, solver_maxit, solver)
        F = problem.F.cpu().numpy()
        loaded = np.where(np.abs(F) > 0.0)[0]
        nodes = np.unique(loaded // 3)
        Ftot = [float(np.sum(np.abs(F[3 * nodes + k]))) if nodes.size else 0.0 for k in range(3)]
        F_total = float(np.sum(np.abs(F[loaded]))) if loaded.size else 0.0
        Uh = U[torch.from_numpy(np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2]))
               .to(problem.dev)].cpu().numpy() if nodes.size else np.zeros(0)
        m = nodes.size
        uavg = [float(np.mean(Uh[k * m:(k + 1) * m])) if m else None for k in range(3)]
        kavg = []
        for k in range(3):
            u = uavg[k]
            kavg.append(float(Ftot[k] / abs(u)) if (u is not None and u != 0.0 and Ftot[k] > 0.0)
                        else None)
        dominant = max(range(3), key=lambda k: Ftot[k])
        return {
            "compliance": float(c),
            "ux_avg_load_patch": uavg[0], "uy_avg_load_patch": uavg[1],
            "uz_avg_load_patch": uavg[2],
            "k_avg_x": kavg[0], "k_avg_y": kavg[1], "k_avg_z": kavg[2], "k_avg": kavg[dominant],
            "F_total": F_total, "F_total_x": Ftot[0], "F_total_y": Ftot[1], "F_total_z": Ftot[2],
            "loaded_x_dof_count": int(m), "loaded_y_dof_count": int(m),
            "loaded_z_dof_count": int(m),
        }
    finally:
        problem.close()


def evaluate_fixed_geometry_compliance(xPhys: np.ndarray, penal: float,
                                       material_params: Optional[Tuple[float, ...]] = None,
                                       elem_size: float = 0.01,
                                       force_field: Optional[np.ndarray] = None,
                                       support_mask: Optional[np.ndarray] = None,
                                       obstacle_mask: Optional[np.ndarray] = None,
                                       protected_zone_mask: Optional[np.ndarray] = None,
                                       use_gpu: bool = False, **kw) -> float:
    """Compliance only (pytopo3d/core/optimizer.py:255-279)."""
    return float(evaluate_fixed_geometry_metrics(
        xPhys=xPhys, penal=penal, material_params=material_params, elem_size=elem_size,
        force_field=force_field, support_mask=support_mask, obstacle_mask=obstacle_mask,
        protected_zone_mask=protected_zone_mask, use_gpu=use_gpu, **kw)["compliance"])


def top3d(
    nelx: int,
    nely: int,
    nelz: int,
    volfrac: float,
    penal: float,
    rmin: float,
    disp_thres: float,
    material_params: Optional[Sequence[float]] = None,
    elem_size: float = 0.01,
    obstacle_mask: Optional[np.ndarray] = None,
    force_field: Optional[np.ndarray] = None,
    support_mask: Optional[np.ndarray] = None,
    tolx: float = 0.01,
    maxloop: int = 2000,
    save_history: bool = False,
    history_frequency: int = 10,
    use_gpu: bool = False,
    protected_zone_mask: Optional[np.ndarray] = None,
    beta_schedule: Sequence[float] = (1.0, 2.0, 4.0, 8.0),
    projection_eta: float = 0.5,
    move: float = 0.2,
    diagnostics_out: Optional[MutableMapping[str, Any]] = None,
    *,
    force_vector: Optional[np.ndarray] = None,
    precision: str = "fp64",
    solver: str = "mg",
    solver_rtol: float = 1e-12,
    solver_maxit: int = 20000,
    solver_options: Optional[dict] = None,
    final_evaluation: bool = True,
    timings: Optional[dict] = None,
    iteration_callback=None,
):
    """Density-filtered SIMP with Heaviside continuation on the GPU.

    Mirrors pytopo3d/core/optimizer.py:283-839: ``maxloop`` is the iteration
    cap of each beta stage; the result is the projected physical density
    ``(nely, nelx, nelz)`` float64, the history dict (or None), the final
    projected compliance and None.
    """
    if min(nelx, nely, nelz) < 1:
        raise ValueError("nelx, nely, and nelz must all be positive")
    expected_shape = (nely, nelx, nelz)
    if not 0.0 < volfrac <= 1.0:
        raise ValueError(f"volfrac must be in (0, 1], got {volfrac}")
    if not np.isfinite(penal) or penal <= 0:
        raise ValueError(f"penal must be positive, got {penal}")
    if not np.isfinite(maxloop) or maxloop < 1:
        raise ValueError(f"maxloop must be at least 1, got {maxloop}")
    if not np.isfinite(tolx) or tolx <= 0:
        raise ValueError(f"tolx must be positive, got {tolx}")
    if not np.isfinite(history_frequency) or history_frequency < 1:
        raise ValueError(f"history_frequency must be at least 1, got {history_frequency}")
    if not np.isfinite(move) or move <= 0:
        raise ValueError(f"move must be positive, got {move}")
    if not 0.0 < projection_eta < 1.0:
        raise ValueError(f"projection_eta must be between 0 and 1, got {projection_eta}")
    beta_schedule = tuple(float(beta) for beta in beta_schedule)
    if not beta_schedule:
        raise ValueError("beta_schedule must contain at least one value")
    if any(not np.isfinite(beta) or beta <= 0.0 for beta in beta_schedule):
        raise ValueError(
            f"all beta_schedule values must be finite and positive, got {beta_schedule}")
    if precision not in _native.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_native.PRECISIONS)}, got {precision!r}")
    if solver not in ("mg", "jacobi"):
        raise ValueError(f"solver must be 'mg' or 'jacobi', got {solver!r}")

    solid, void, free = protected_masks(nelx, nely, nelz, obstacle_mask, protected_zone_mask,
                                        force_field, support_mask, force_vector, logger)
    design_nele = int(np.sum(free))
    if design_nele == 0:
        raise ValueError("optimization requires at least one free design element")
    target_free_volume = float(volfrac * design_nele)
    logger.debug("free design elements: %d/%d", design_nele, nelx * nely * nelz)

    t_setup = time.perf_counter()
    problem = _make_problem(nelx, nely, nelz, rmin, material_params, elem_size, force_field,
                            support_mask, solid, void, force_vector, solver_options,
                            obstacle_mask=obstacle_mask, protected_zone_mask=protected_zone_mask,
                            device_flags=True)
    try:
        plan = problem.plan
        dev = problem.dev
        logger.info("Linear solver: matrix-free H8 %s-PCG (%s) on %s", solver, precision,
                    torch.cuda.get_device_name(dev))
        history = ({"density_history": [], "iteration_history": [], "compliance_history": [],
                    "beta_history": []} if save_history else None)
        solve_kw = dict(precision=precision, rtol=solver_rtol, maxit=solver_maxit, solver=solver)

        x0 = np.full(expected_shape, float(volfrac))
        x0[solid] = 1.0
        x0[void] = 0.0
        st = DesignState(problem, x0)
        if save_history:
            st.snapshots = HistorySnapshots(expected_shape)
        t_setup = time.perf_counter() - t_setup

        tm = {"setup_s": t_setup, "solve_s": 0.0, "sens_s": 0.0, "oc_s": 0.0, "iter_s": 0.0,
              "solve_iters": [], "oc_steps": []}
        total_loop = 0
        c = float("nan")
        stage_summaries = []
        eta = projection_eta
        for beta in beta_schedule:
            logger.info("Starting projection stage: beta=%g", beta)
            change = float("inf")
            stage_loop = 0
            c_prev = float("nan")
            plan.filter_project(st.x, beta, eta, st.rf, st.rho, st.drho)
            vmin, vmax = plan.volume_bounds(beta, eta)
            target_within = vmin - 1.0e-10 <= volfrac <= vmax + 1.0e-10
            if volfrac < vmin - 1.0e-10:
                raise ValueError(
                    "projected free-volume target is infeasible at "
                    f"beta={beta:g}: target={volfrac:.6f}, minimum achievable="
                    f"{vmin:.6f}. Protected-solid density "
                    "spreads into the free region through the density filter.")
            if volfrac > vmax + 1.0e-10:
                logger.warning(
                    "Projected free-volume target %.6f exceeds the maximum "
                    "achievable %.6f at beta=%g; the volume upper bound is "
                    "inactive at the all-solid free design.", volfrac, vmax, beta)
            vol_frac = float("nan")
            gray = float("nan")
            while change > tolx and stage_loop < maxloop:
                stage_loop += 1
                total_loop += 1
                t0 = time.perf_counter()
                snapshot = save_history and ((total_loop - 1) % history_frequency == 0
                                             or stage_loop == maxloop)
                r = design_step(problem, st, penal, beta, eta, move, target_free_volume, solve_kw)
                c = r["c"]
                change, vol_frac, gray = r["change"], r["vol"], r["gray"]
                tm["solve_s"] += r["t_solve"]
                tm["sens_s"] += r["t_sens"]
                tm["oc_s"] += r["t_oc"]
                tm["solve_iters"].append(r["solve_iters"])
                tm["oc_steps"].append(r["oc_steps"])
                if iteration_callback is not None:
                    iteration_callback({"beta": beta, "iter": stage_loop, "total": total_loop,
                                        "c": c, "change": change, "vol": vol_frac, "gray": gray,
                                        "x": st.x, "rho": st.rho_alt})
                c_delta = c - c_prev if np.isfinite(c_prev) else float("nan")
                c_prev = c
                elapsed = time.perf_counter() - t0
                tm["iter_s"] += elapsed
                if save_history and (snapshot or change <= tolx):
                    # st.rho_alt holds the density this step solved with; its
                    # copy to pinned host memory overlaps the next solve
                    st.snapshots.take(st.rho_alt)
                    history["density_history"].append(len(st.snapshots.host) - 1)
                    history["iteration_history"].append(total_loop - 1)
                    history["compliance_history"].append(c)
                    history["beta_history"].append(beta)
                logger.info(
                    "beta=%4.1f iter=%4d total=%4d C=%.6e dC=%.3e "
                    "vol=%.5f gray=%.5f change=%.5e time=%.2fs",
                    beta, stage_loop, total_loop, c, c_delta, vol_frac, gray, change, elapsed)
            stage_converged = change <= tolx
            if not stage_converged:
                logger.warning(
                    "Projection stage beta=%g reached maxloop=%d without "
                    "converging: final change %.6e exceeds tolx %.6e.", beta, maxloop, change, tolx)
            stage_summaries.append({
                "beta": float(beta),
                "iterations": int(stage_loop),
                "converged": bool(stage_converged),
                "physical_density_fraction": float(vol_frac),
                "volume_fraction_error": float(vol_frac) - float(volfrac),
                "gray_fraction_005_095": float(gray),
                "final_change": float(change),
                "minimum_achievable_physical_fraction": vmin,
                "maximum_achievable_physical_fraction": vmax,
                "target_within_achievable_range": target_within,
            })

        if save_history:
            snaps = st.snapshots.arrays()
            history["density_history"] = [snaps[i] for i in history["density_history"]]
        final_beta = beta_schedule[-1]
        plan.filter_project(st.x, final_beta, eta, st.rf, st.rho, st.drho)
        final_x = _to_host_elem(st.x, expected_shape)
        final_rf = _to_host_elem(st.rf, expected_shape)
        final_xPhys = _to_host_elem(st.rho, expected_shape)
        if not final_evaluation:
            if timings is not None:
                timings.update(tm, counters=plan.counters())
            return final_xPhys, history, float("nan"), None

        ts = time.perf_counter()
        final_compliance, _ = _fixed_geometry(problem, final_xPhys, penal, precision, solver_rtol,
                                              solver_maxit, solver)
        if save_history:
            history["density_history"].append(final_xPhys.copy())
            history["iteration_history"].append(total_loop)
            history["compliance_history"].append(float(final_compliance))
            history["beta_history"].append(float(final_beta))
        rho_binary = (final_xPhys >= 0.5).astype(float)
        rho_binary[solid] = 1.0
        rho_binary[void] = 0.0
        binary_compliance, _ = _fixed_geometry(problem, rho_binary, penal, precision, solver_rtol,
                                               solver_maxit, solver)
        tm["final_eval_s"] = time.perf_counter() - ts
        binary_delta = ((binary_compliance - final_compliance) / final_compliance
                        if final_compliance != 0.0 else None)
        free_values = final_xPhys[free]
        gray_mask = (free_values > 0.05) & (free_values < 0.95)
        projection_metrics = {
            "projection_enabled": True,
            "projection_beta": float(final_beta),
            "projection_beta_schedule": [float(b) for b in beta_schedule],
            "projection_eta": float(projection_eta),
            "penal": float(penal),
            "design_variable_fraction": float(np.mean(final_x[free])),
            "filtered_density_fraction": float(np.mean(final_rf[free])),
            "physical_density_fraction": float(np.mean(free_values)),
            "gray_fraction_005_095": float(np.mean(gray_mask.astype(float))),
            "projected_compliance": float(final_compliance),
            "binary_compliance": float(binary_compliance),
            "binary_compliance_delta": None if binary_delta is None else float(binary_delta),
            "binary_threshold": 0.5,
            "target_free_volume_fraction": float(volfrac),
            "physical_volume_fraction_error": float(np.mean(free_values) - volfrac),
            "projection_total_iterations": int(total_loop),
            "projection_stage_summaries": stage_summaries,
            "projection_converged": all(s["converged"] for s in stage_summaries),
            "projection_target_within_achievable_range": all(
                s["target_within_achievable_range"] for s in stage_summaries),
        }
        logger.info(
            "Final projection: beta=%.1f eta=%.3f vol=%.5f gray=%.5f "
            "C_projected=%.6e C_binary=%.6e binary_delta=%s", final_beta, projection_eta,
            projection_metrics["physical_density_fraction"], projection_metrics["gray_fraction_005_095"],
            final_compliance, binary_compliance,
            "n/a" if binary_delta is None else f"{binary_delta:.6e}")
        if diagnostics_out is not None:
            diagnostics_out.clear()
            diagnostics_out.update(projection_metrics)
        if timings is not None:
            timings.update(tm, counters=plan.counters())
        return final_xPhys, history, float(final_compliance), None
    finally:
        problem.close()
