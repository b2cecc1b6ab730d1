#!/usr/bin/env python3
"""Design iterations/s of the SIMP hot path at 256x128x128 H8 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (synthetic, no data files): the 256x128x128 unit-cube H8 cantilever
-- all DOFs fixed on x = 0, a total unit force in -y spread over the nodes of
the edge x = nelx, y = 0 -- volfrac 0.3, penal 3, rmin 3, eta 0.5, move 0.2,
x0 = 0.3 with the protected strips imposed, beta = 1 (the first continuation
stage).  A step is one full design iteration on the device: linear solve
(matrix-free multigrid-PCG), energies + chain rule, OC bisection, re-filtering
and the free-volume statistics, with the host reading the per-iteration
scalars exactly as top3d does.

The headline value is the mixed-precision path (FP64 flexible conjugate
gradients whose preconditioner is an FP32 multigrid V-cycle, to a relative
residual of 1e-8).  The same
steps are then timed on the FP64 path at the same tolerance, and the
compliance / sensitivity error of the mixed solve against an FP64 solve of the
same design is reported.  The roofline object is for the dominant kernel
class, timed with HIP events on the stream the library launches on.  The CPU
baseline times the CPU oracle (assembled K + SuperLU, the reference's CPU
path without PyPardiso) on a bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "design iters/s at 256x128x128 H8: mixed vs FP64 GPU, c err <=1e-6 vs FP64"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
# arithmetic operations the Walsh-Hadamard sector form executes per element and
# element layer (24 + 21 + 45 + 21 + 21 + 12 + 24), against 1152 flops of the
# dense 24x24 product the "dense_equiv" figures count
SECTOR_OPS = 168
# VALU lane operations per second without packed FP32: 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz (an FMA counts as one operation)
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9


def _parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nelx", type=int, default=256)
    ap.add_argument("--nely", type=int, default=128)
    ap.add_argument("--nelz", type=int, default=128)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="CPU seconds per sample size of the CPU baseline (two sizes)")
    ap.add_argument("--no-operator-rates", action="store_true",
                    help="skip the K u applications/s at every benchmark grid")
    ap.add_argument("--config4", action="store_true",
                    help="run only the 320x160x160 combined configuration (volfrac 0.2, rmin 4, "
                         "beta 1, 2, 4, 8 x 40 iterations) on both paths")
    ap.add_argument("--no-config4", action="store_true",
                    help="skip the 320x160x160 combined configuration in the default run")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--smooth-degree", type=int, default=0,
                    help="Chebyshev degree on the two finest levels (0: default)")
    ap.add_argument("--coarse-degree", type=int, default=0,
                    help="Chebyshev degree on the coarser levels (0: default)")
    ap.add_argument("--mid-degree", type=int, default=0,
                    help="Chebyshev degree on the coarse levels of more than 32768 elements "
                         "(0: default)")
    ap.add_argument("--mid-cheb-lo", type=float, default=0.0,
                    help="smoothing interval low end on those levels (0: default)")
    ap.add_argument("--cheb-lo", type=float, default=0.0, help="smoothing interval low end (0: default)")
    ap.add_argument("--coarse-cheb-lo", type=float, default=0.0,
                    help="smoothing interval low end on the coarser levels (0: default)")
    ap.add_argument("--cheb-hi", type=float, default=0.0,
                    help="smoothing interval high end, fraction of the lambda_max bound (0: default)")
    ap.add_argument("--galerkin", type=int, default=0,
                    help="third multigrid level: 0 default (rediscretised), 1 Galerkin, 2 rediscretised")
    ap.add_argument("--operator-kind", type=int, default=0,
                    help="0: element/sector kernel (default), 1: dense node kernel")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="only launch the hot kernels a fixed number of times (for rocprofv3 "
                         "counter runs) plus a 1 GiB reference copy, then exit")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def _log(args, *msg):
    print("[bench]", *msg, file=sys.stderr, flush=True)


def _setup(args):
    import torch
    from pytopo3d_amd.optimizer import DesignState, SIMPProblem
    from pytopo3d_amd.problem import (edge_load_force_vector, element_stiffness,
                                      protected_masks, support_node_flags)
    nelx, nely, nelz = args.nelx, args.nely, args.nelz
    F = edge_load_force_vector(nelx, nely, nelz)
    solid, void, free = protected_masks(nelx, nely, nelz, None, None, None, None, F)
    KE = element_stiffness(None, 1.0)
    fixed = support_node_flags(nelx, nely, nelz, None)
    problem = SIMPProblem(nelx, nely, nelz, 3.0, KE, F, fixed, solid, void,
                          solver_options=_solver_options(args))
    x0 = np.full((nely, nelx, nelz), 0.3)
    x0[solid] = 1.0
    x0[void] = 0.0
    target = 0.3 * int(free.sum())
    return problem, x0, target, DesignState, torch


def _solver_options(args):
    """tk_options fields set on the command line (options left at 0 keep the
    library defaults, and are reported as such)."""
    opts = {"operator_kind": args.operator_kind, "smooth_degree": args.smooth_degree,
            "coarse_degree": args.coarse_degree, "cheb_lo": args.cheb_lo, "cheb_hi": args.cheb_hi,
            "coarse_cheb_lo": args.coarse_cheb_lo, "galerkin": args.galerkin,
            "mid_degree": args.mid_degree, "mid_cheb_lo": args.mid_cheb_lo}
    return {k: v for k, v in opts.items() if v}


def _run_path(problem, x0, target, DesignState, torch, precision, args):
    """W untimed + K timed design iterations; returns (seconds, per-step records, state)."""
    from pytopo3d_amd.optimizer import design_step
    st = DesignState(problem, x0)
    problem.have_U = False
    beta, eta = 1.0, 0.5
    problem.plan.filter_project(st.x, beta, eta, st.rf, st.rho, st.drho)
    kw = dict(precision=precision, rtol=args.rtol, maxit=20000, solver="mg")
    for i in range(args.warmup):
        r = design_step(problem, st, 3.0, beta, eta, 0.2, target, kw)
        _log(args, f"{precision} warmup {i}: c={r['c']:.6e} iters={r['solve_iters']} "
                   f"oc={r['oc_steps']} solve={r['t_solve']*1e3:.1f}ms oc={r['t_oc']*1e3:.1f}ms")
    torch.cuda.synchronize()
    recs = []
    c0 = problem.plan.counters()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = design_step(problem, st, 3.0, beta, eta, 0.2, target, kw)
        recs.append(r)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    c1 = problem.plan.counters()
    work = {k: c1[k] - c0[k] for k in c1}
    for i, r in enumerate(recs):
        _log(args, f"{precision} step {i}: c={r['c']:.6e} iters={r['solve_iters']} oc={r['oc_steps']} "
                   f"solve={r['t_solve']*1e3:.1f}ms sens={r['t_sens']*1e3:.1f}ms oc={r['t_oc']*1e3:.1f}ms")
    return dt, recs, st, work


def _time_kernels(problem, st, torch, args):
    """Average launch duration (HIP events on the library's stream) of the
    kernel classes, with their algorithmic bytes and flops per launch."""
    plan = problem.plan
    nele, ndof = plan.nele, plan.ndof
    dev = problem.dev
    reps = args.kernel_reps
    out = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    # operator, FP32 and FP64, on the state tk_prepare builds for the design
    u32 = torch.randn(ndof, dtype=torch.float32, device=dev)
    y32 = torch.empty_like(u32)
    u64 = u32.double()
    y64 = torch.empty_like(u64)
    mf = ", true" if plan.options["operator_kind"] == 2 else ""
    plan.prepare(st.rho, 3.0, precision="mixed")
    t = timed(lambda: plan.apply_K(None, 3.0, u32, y32, precision="mixed"))
    out["operator_fp32"] = {"s": t, "bytes": 4 * (2 * ndof + nele), "flops": 1152 * nele,
                            "ops": SECTOR_OPS * nele, "kernel": f"k_apply_el_w3<float, 0{mf}>"}
    plan.prepare(st.rho, 3.0, precision="fp64")
    t = timed(lambda: plan.apply_K(None, 3.0, u64, y64, precision="fp64"))
    out["operator_fp64"] = {"s": t, "bytes": 8 * (2 * ndof + nele), "flops": 1152 * nele,
                            "ops": SECTOR_OPS * nele, "kernel": "k_apply_el_w3<double, 0>"}
    # the dominant kernel of the mixed solve: the FP32 Chebyshev smoothing step
    plan.prepare(st.rho, 3.0, precision="mixed")
    b32 = torch.randn(ndof, dtype=torch.float32, device=dev)
    d32 = torch.zeros_like(b32)
    xo32 = torch.empty_like(b32)
    t = timed(lambda: plan.smoother_step(b32, u32, xo32, d32, 0.5, 0.1, precision="mixed"))
    # reads x, b, d and the element scales (the inverse diagonal is formed from
    # the scales), writes x' and d
    out["smoother_fp32"] = {"s": t, "bytes": 4 * (5 * ndof + nele), "flops": 1152 * nele,
                            "ops": SECTOR_OPS * nele, "kernel": f"k_apply_el<float, 3{mf}>"}
    rf, rho, d = problem.new_elem(), problem.new_elem(), problem.new_elem()
    t = timed(lambda: plan.filter_project(st.x, 1.0, 0.5, rf, rho, d))
    out["filter_project"] = {"s": t, "bytes": nele * (8 + 8 + 1 + 24),
                             "flops": 2 * 93 * nele, "ops": 2 * 93 * nele,
                             "kernel": "k_filter_tiled<0, 2>"}
    return out


# Benchmark configurations whose operator applications/s are reported
# (plain 60x30x20, 128x64x64, 96x48x48, the headline 256x128x128, 320x160x160).
OPERATOR_GRIDS = [(60, 30, 20), (128, 64, 64), (96, 48, 48), (256, 128, 128), (320, 160, 160)]


def _operator_rates(args):
    """K u applications/s (FP64 and FP32, one full-grid launch each, HIP
    events) on a seeded uniform [0.01, 1] design at each benchmark grid, with
    the HBM fraction of the algorithmic bytes 8 (2 ndof + nele) / 4 (...)."""
    import torch
    from pytopo3d_amd import _native
    from pytopo3d_amd.problem import element_flags, element_stiffness, support_node_flags
    KE = element_stiffness(None, 1.0)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = {}
    for nelx, nely, nelz in OPERATOR_GRIDS:
        plan = _native.Plan(nelx, nely, nelz, KE, 1.5,
                            {"operator_kind": args.operator_kind} if args.operator_kind else None)
        z = np.zeros((nely, nelx, nelz), bool)
        plan.set_masks(torch.from_numpy(element_flags(z, z)).to(dev),
                       torch.from_numpy(support_node_flags(nelx, nely, nelz)).to(dev))
        rho = torch.from_numpy(np.random.default_rng(7).uniform(0.01, 1.0, plan.nele)).to(dev)
        row = {}
        for prec, dt, nb in (("fp64", torch.float64, 8), ("mixed", torch.float32, 4)):
            u = torch.randn(plan.ndof, dtype=dt, device=dev)
            y = torch.empty_like(u)
            plan.prepare(rho, 3.0, precision=prec)
            plan.apply_K(None, 3.0, u, y, precision=prec)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.kernel_reps):
                plan.apply_K(None, 3.0, u, y, precision=prec)
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / 1e3 / args.kernel_reps
            key = "fp64" if prec == "fp64" else "fp32"
            row[key] = {"applies_per_s": 1.0 / t, "us": t * 1e6,
                        "hbm_frac": nb * (2 * plan.ndof + plan.nele) / t / 1e9 / HBM_PEAK_GBS}
        out[f"{nelx}x{nely}x{nelz}"] = row
        plan.close()
        torch.cuda.empty_cache()
    return out


def _kernel_probe(args):
    """Fixed launches for counter collection: 5 x FP32 operator, 5 x FP32
    smoothing step, 5 x FP64 operator, 5 x filter+projection and 5 x a 1 GiB
    device copy whose FETCH_SIZE / WRITE_SIZE calibrate the counters.  No solve
    runs, so every launch of the operator kernels here is a full fine-grid
    launch and their rocprofv3 average is the per-launch time."""
    problem, x0, target, DesignState, torch = _setup(args)
    plan = problem.plan
    st = DesignState(problem, x0)
    plan.filter_project(st.x, 1.0, 0.5, st.rf, st.rho, st.drho)
    ndof = plan.ndof
    u32 = torch.randn(ndof, dtype=torch.float32, device=problem.dev)
    y32 = torch.empty_like(u32)
    u64, y64 = u32.double(), torch.empty(ndof, dtype=torch.float64, device=problem.dev)
    plan.prepare(st.rho, 3.0, precision="mixed")
    torch.cuda.synchronize()
    for _ in range(5):
        plan.apply_K(None, 3.0, u32, y32, precision="mixed")
    b32, d32, xo32 = torch.randn_like(u32), torch.zeros_like(u32), torch.empty_like(u32)
    for _ in range(5):
        plan.smoother_step(b32, u32, xo32, d32, 0.5, 0.1, precision="mixed")
    plan.prepare(st.rho, 3.0, precision="fp64")
    torch.cuda.synchronize()
    for _ in range(5):
        plan.apply_K(None, 3.0, u64, y64, precision="fp64")
    rf, rho, d = problem.new_elem(), problem.new_elem(), problem.new_elem()
    for _ in range(5):
        plan.filter_project(st.x, 1.0, 0.5, rf, rho, d)
    a = torch.ones(1 << 27, dtype=torch.float64, device=problem.dev)   # 1 GiB
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "ndof": ndof, "nele": plan.nele,
                      "copy_bytes": a.numel() * 8}), flush=True)
    problem.close()


def _traffic(kernel_prefix):
    """HBM bytes per launch from the committed counter summary (None if absent)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t["kernels"][kernel_prefix]["hbm_bytes_per_launch"]
    except Exception:
        return None


def _cpu_rate(nelx, nely, nelz, seconds):
    """Design iterations/s of the oracle (COO->CSR assembly, SuperLU, sparse
    filter, OC bisection) on one host core at one grid, beta 1."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    F = oracle.edge_load_force_vector(nelx, nely, nelz)
    solid, void, free = oracle.boundary_protection(nelx, nely, nelz, None, None, None, None, F)
    KE = oracle.element_stiffness(None, 1.0)
    freedofs, _ = oracle.support_dofs(nelx, nely, nelz)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, 3.0)
    edof = oracle.element_dofs(nelx, nely, nelz)
    shape = (nely, nelx, nelz)
    x = np.full(shape, 0.3)
    x[solid], x[void] = 1.0, 0.0
    target = 0.3 * free.sum()
    _, rho, d = oracle.physical_density(x, H, Hs, 1.0, 0.5, solid, void)
    n = 0
    t0 = time.perf_counter()
    while True:
        U = oracle.solve_displacement(rho, 3.0, KE, F, freedofs, nelx, nely, nelz)
        ce = oracle.element_energy(U, edof, KE).reshape(shape, order="F")
        dc = -3.0 * (1 - 1e-9) * rho ** 2.0 * ce
        dcdx = oracle.filter_chain_rule(dc, d, H, Hs)
        dvdx = oracle.filter_chain_rule(free.astype(float), d, H, Hs)
        x, _ = oracle.oc_projected(x, dcdx, dvdx, 0.2, target, H, Hs, 1.0, 0.5, free, solid, void)
        _, rho, d = oracle.physical_density(x, H, Hs, 1.0, 0.5, solid, void)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return n, time.perf_counter() - t0


def _cpu_baseline(args):
    """The oracle's design iteration (the reference's CPU path with SuperLU,
    PyPardiso not being installed) measured on this host at two grids; the
    cost exponent fitted to the two measurements carries the rate to the
    benchmark grid, where the assembled problem alone would need ~77 GB."""
    pts = []
    for dims in ((20, 10, 7), (30, 15, 10)):
        n, dt = _cpu_rate(*dims, args.cpu_seconds)
        pts.append((dims, n, dt))
    (d1, n1, t1), (d2, n2, t2) = pts
    e1, e2 = np.prod(d1), np.prod(d2)
    s1, s2 = t1 / n1, t2 / n2                          # seconds per design iteration
    alpha = float(np.log(s2 / s1) / np.log(e2 / e1))
    target = args.nelx * args.nely * args.nelz
    value = 1.0 / (s2 * (target / e2) ** alpha)
    return {
        "value": value,
        "unit": "design iters/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle design iterations (assembled K + SuperLU, single core): "
                   f"{n1} at {'x'.join(map(str, d1))} in {t1:.1f} s ({1 / s1:.3f} iters/s), "
                   f"{n2} at {'x'.join(map(str, d2))} in {t2:.1f} s ({1 / s2:.3f} iters/s); "
                   f"time per iteration grows as elements^{alpha:.2f} between them, which "
                   f"carries the rate to {args.nelx}x{args.nely}x{args.nelz}; an upper bound "
                   f"on the CPU rate: one iteration at 60x30x20 took 447 s on one core "
                   f"(measured once, outside this run), 10x what this exponent predicts, "
                   f"as the direct solve's fill grows faster on larger grids"),
        "measured": [{"grid": list(d), "iters": n, "seconds": t} for d, n, t in pts],
        "exponent": alpha,
    }


def _config4(args):
    """320x160x160 combined configuration: volfrac 0.2, penal 3, rmin 4, beta 1,
    2, 4, 8 with 40 design iterations each, edge-load cantilever, through
    top3d on the mixed path (FP32 multigrid inside FP64 flexible CG, rtol 1e-8)
    and the FP64 path; design iterations/s, final compliance of each, peak
    device memory."""
    import torch
    from pytopo3d_amd import edge_load_force_vector, top3d
    nelx, nely, nelz = 320, 160, 160
    F = edge_load_force_vector(nelx, nely, nelz)
    kw = dict(volfrac=0.2, penal=3.0, rmin=4.0, disp_thres=0.5, force_vector=F, maxloop=40,
              tolx=1e-12, beta_schedule=(1.0, 2.0, 4.0, 8.0), projection_eta=0.5, move=0.2,
              solver_rtol=args.rtol, solver_options=_solver_options(args))
    out = {"config": "320x160x160, volfrac 0.2, rmin 4, beta 1,2,4,8 x 40 iterations"}
    for prec in ("mixed", "fp64"):
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        free0, total = torch.cuda.mem_get_info()
        low = [free0]
        tm = {}

        def cb(d):
            if d["total"] % 20 == 0:
                low.append(torch.cuda.mem_get_info()[0])
                _log(args, f"config4 {prec} iteration {d['total']}: c={d['c']:.6e}")

        t0 = time.perf_counter()
        _, _, c, _ = top3d(nelx, nely, nelz, precision=prec, timings=tm, iteration_callback=cb,
                           **kw)
        dt = time.perf_counter() - t0
        iters = len(tm["solve_iters"])
        out[prec] = {"design_iters": iters, "loop_s": tm["iter_s"],
                     "design_iters_per_s": iters / tm["iter_s"], "wall_s": dt,
                     "final_compliance": c, "cg_iters_mean": float(np.mean(tm["solve_iters"])),
                     "peak_device_gb": (free0 - min(low)) / 1e9}
    out["c_rel_diff_mixed_vs_fp64"] = abs(out["mixed"]["final_compliance"] -
                                          out["fp64"]["final_compliance"]) / abs(
        out["fp64"]["final_compliance"])
    out["speedup"] = out["mixed"]["design_iters_per_s"] / out["fp64"]["design_iters_per_s"]
    # the dominant kernel at this size: the FP32 smoothing step on a seeded
    # uniform [0.01, 1] design, HIP events, algorithmic bytes 4 (5 ndof + nele)
    from pytopo3d_amd.optimizer import SIMPProblem
    from pytopo3d_amd.problem import element_stiffness, protected_masks, support_node_flags
    solid, void, _ = protected_masks(nelx, nely, nelz, None, None, None, None, F)
    pr = SIMPProblem(nelx, nely, nelz, 4.0, element_stiffness(None, 1.0), F,
                     support_node_flags(nelx, nely, nelz), solid, void)
    plan = pr.plan
    rho = torch.from_numpy(np.random.default_rng(4).uniform(0.01, 1.0, plan.nele)).to(pr.dev)
    plan.prepare(rho, 3.0, precision="mixed")
    b, x = (torch.randn(plan.ndof, dtype=torch.float32, device=pr.dev) for _ in range(2))
    d, xo = torch.zeros_like(b), torch.empty_like(b)
    plan.smoother_step(b, x, xo, d, 0.5, 0.1, precision="mixed")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.kernel_reps):
        plan.smoother_step(b, x, xo, d, 0.5, 0.1, precision="mixed")
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / args.kernel_reps
    nbytes = 4 * (5 * plan.ndof + plan.nele)
    out["smoother_fp32"] = {"us": t * 1e6, "GBps": nbytes / t / 1e9,
                            "hbm_frac": nbytes / t / 1e9 / HBM_PEAK_GBS}
    pr.close()
    return out


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.kernel_probe:
        _kernel_probe(args)
        return
    if args.config4:
        print(json.dumps(_config4(args)), flush=True)
        return
    problem, x0, target, DesignState, torch = _setup(args)
    plan = problem.plan
    _log(args, "levels", plan.levels(), "nfree", problem.nfree)

    t_mixed, recs_m, st_m, work_m = _run_path(problem, x0, target, DesignState, torch, "mixed", args)
    value = args.steps / t_mixed
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "design iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.nelx}x{args.nely}x{args.nelz} H8 cantilever (edge load), "
                         "design iterations at beta=1, FP64 flexible CG with an FP32 "
                         f"multigrid correction, to rtol {args.rtol:g}"),
            "nelx": args.nelx, "nely": args.nely, "nelz": args.nelz,
            "volfrac": 0.3, "penal": 3.0, "rmin": 3.0, "beta": 1.0, "eta": 0.5, "move": 0.2,
            "solver_rtol": args.rtol, "levels": plan.levels(),
            "smooth_degree": plan.options["smooth_degree"],
            "coarse_degree": plan.options["coarse_degree"], "cheb_lo": plan.options["cheb_lo"],
            "coarse_cheb_lo": plan.options["coarse_cheb_lo"],
            "mid_degree": plan.options["mid_degree"], "mid_cheb_lo": plan.options["mid_cheb_lo"],
        },
        "solve_iters_mean": float(np.mean([r["solve_iters"] for r in recs_m])),
        "oc_steps_mean": float(np.mean([r["oc_steps"] for r in recs_m])),
        "oc_evals_mean": float(np.mean([r["oc_evals"] for r in recs_m])),
        "work_timed_mixed": work_m,
        "phase_ms_mixed": {k: 1e3 * float(np.mean([r[k] for r in recs_m]))
                           for k in ("t_solve", "t_sens", "t_oc")},
    }
    if not args.no_fp64:
        t64, recs_d, st_d, work_d = _run_path(problem, x0, target, DesignState, torch, "fp64", args)
        result["fp64_value"] = args.steps / t64
        result["speedup_vs_fp64"] = value / result["fp64_value"]
        result["work_timed_fp64"] = work_d
        result["phase_ms_fp64"] = {k: 1e3 * float(np.mean([r[k] for r in recs_d]))
                                   for k in ("t_solve", "t_sens", "t_oc")}
        # compliance / sensitivity error of the mixed solve on the mixed run's design
        rho = st_m.rho
        Ua = torch.zeros(plan.ndof, dtype=torch.float64, device=problem.dev)
        Ub = torch.zeros_like(Ua)
        plan.solve(rho, 3.0, problem.F, Ua, rtol=1e-11, precision="fp64")
        plan.solve(rho, 3.0, problem.F, Ub, rtol=args.rtol, precision="mixed")
        ga, gb = problem.new_elem(), problem.new_elem()
        ca = plan.energy(Ua, rho, 3.0, drho=st_m.drho, dc_dx=ga)
        cb = plan.energy(Ub, rho, 3.0, drho=st_m.drho, dc_dx=gb)
        result["c_rel_err_vs_fp64"] = abs(cb - ca) / abs(ca)
        result["dc_rel_err_vs_fp64"] = float((gb - ga).abs().max() / ga.abs().max())
        result["c_first_step_rel_diff"] = abs(recs_m[0]["c"] - recs_d[0]["c"]) / abs(recs_d[0]["c"])
    kern = _time_kernels(problem, st_m, torch, args)
    dom = kern["smoother_fp32"]
    achieved = dom["bytes"] / dom["s"] / 1e9
    result["roofline"] = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic(dom["kernel"]),
        "kernel": dom["kernel"],
        "launch_us": dom["s"] * 1e6,
        # executed sector-form operations per second as a fraction of the VALU
        # lane-operation rate, and the dense-equivalent 24x24 rate (not executed)
        "executed_ops_frac_valu": dom["ops"] / dom["s"] / VALU_LANE_OPS,
        "dense_equiv_flop_frac_fp32": dom["flops"] / dom["s"] / 1e12 / FP32_PEAK_TFLOPS,
    }
    result["kernels"] = {k: {"kernel": v["kernel"], "us": v["s"] * 1e6,
                             "GBps": v["bytes"] / v["s"] / 1e9,
                             "hbm_frac": v["bytes"] / v["s"] / 1e9 / HBM_PEAK_GBS,
                             "executed_Tops": v["ops"] / v["s"] / 1e12,
                             "dense_equiv_TFLOPs": v["flops"] / v["s"] / 1e12}
                         for k, v in kern.items()}
    if not args.no_operator_rates:
        result["operator_applies"] = _operator_rates(args)
    problem.close()
    if not args.no_config4:
        result["config4"] = _config4(args)
    if not args.no_cpu:
        result["cpu_baseline"] = _cpu_baseline(args)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
