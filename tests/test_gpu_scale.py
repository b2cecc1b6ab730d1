"""Parity at the full sizes of the benchmark configurations (FP64 unless stated).

* 128x64x64 (the plain FP64 configuration) and 256x128x128 (the headline
  configuration), the cantilever supports (x = 0 fixed), seeded uniform
  [0.01, 1] densities (seed 1 and 2, element order): the FP64 and FP32 K v
  against the oracle's element-by-element product (gather, KE product,
  scatter; no assembly), and diag(K).
* 256x128x128: the mixed-precision solve (FP32 multigrid inside FP64 flexible
  CG, rtol 1e-8) against the FP64 solve (rtol 1e-11) of the same design, with
  the errors of u, c and dc/drho stated; the FP64 solution's residual measured
  with the oracle's operator; then four design iterations of the cantilever
  (edge load, beta 1) on both paths.
"""

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import dev, elem_vec, make_plan, note, to_dev
from pytopo3d_amd import SIMPProblem, edge_load_force_vector
from pytopo3d_amd.optimizer import DesignState, design_step
from pytopo3d_amd.problem import element_stiffness, protected_masks, support_node_flags

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("nelx,nely,nelz,seed", [(128, 64, 64, 1), (256, 128, 128, 2)])
def test_operator_matches_element_by_element_oracle(nelx, nely, nelz, seed):
    note("operator vs oracle", nelx, nely, nelz)
    P, KE = make_plan(nelx, nely, nelz, 1.5)
    nele = nelx * nely * nelz
    rho = np.random.default_rng(seed).uniform(0.01, 1.0, nele)
    u = np.random.default_rng(seed + 100).standard_normal(P.ndof)
    _, fixed = oracle.support_dofs(nelx, nely, nelz)
    want = oracle.apply_stiffness_ebe(u, rho, 3.0, KE, nelx, nely, nelz, fixed)
    rd, ud = to_dev(rho), to_dev(u)
    y = torch.empty_like(ud)
    P.apply_K(rd, 3.0, ud, y)
    e64 = _rel(y.cpu().numpy(), want)
    y32 = torch.empty(P.ndof, dtype=torch.float32, device=dev())
    P.apply_K(rd, 3.0, ud.float(), y32, precision="mixed")
    e32 = _rel(y32.double().cpu().numpy(), want)
    note("  K v rel err FP64", e64, "FP32", e32)
    assert e64 <= 1e-13
    assert e32 <= 2e-6
    dg = torch.empty_like(ud)
    P.diag_K(rd, 3.0, dg)
    dwant = oracle.diag_stiffness_ebe(rho, 3.0, KE, nelx, nely, nelz)
    dwant[fixed] = 1.0
    np.testing.assert_allclose(dg.cpu().numpy(), dwant, rtol=1e-14)
    P.close()


def _headline_problem():
    nelx, nely, nelz = 256, 128, 128
    F = edge_load_force_vector(nelx, nely, nelz)
    solid, void, free = protected_masks(nelx, nely, nelz, None, None, None, None, F)
    KE = element_stiffness(None, 1.0)
    fixed = support_node_flags(nelx, nely, nelz, None)
    return SIMPProblem(nelx, nely, nelz, 3.0, KE, F, fixed, solid, void), F, solid, void, free


def test_mixed_solve_against_fp64_256x128x128():
    note("mixed vs fp64 solve 256x128x128")
    pr, F, solid, void, free = _headline_problem()
    nelx, nely, nelz = pr.nelx, pr.nely, pr.nelz
    rho = to_dev(np.random.default_rng(2).uniform(0.01, 1.0, pr.plan.nele))
    U64 = pr.solve(rho, 3.0, precision="fp64", rtol=1e-11, warm=False, out=torch.zeros_like(pr.F))
    it64 = pr.last_iters
    U32 = pr.solve(rho, 3.0, precision="mixed", rtol=1e-8, warm=False, out=torch.zeros_like(pr.F))
    it32 = pr.last_iters
    n = pr.plan.nele
    dc64, dc32 = (torch.empty(n, dtype=torch.float64, device=dev()) for _ in range(2))
    c64 = pr.plan.energy(U64, rho, 3.0, dc_drho=dc64)
    c32 = pr.plan.energy(U32, rho, 3.0, dc_drho=dc32)
    u64 = U64.cpu().numpy()
    eu = _rel(U32.cpu().numpy(), u64)
    ec = abs(c32 - c64) / abs(c64)
    edc = _rel(dc32.cpu().numpy(), dc64.cpu().numpy())
    note("  iterations fp64", it64, "mixed", it32, "| rel err u", eu, "c", ec, "dc", edc)
    # the FP64 solution's residual with the oracle's element-by-element operator
    freedofs, fixed = oracle.support_dofs(nelx, nely, nelz)
    KE = oracle.element_stiffness(None, 1.0)
    ku = oracle.apply_stiffness_ebe(u64, rho.cpu().numpy(), 3.0, KE, nelx, nely, nelz, fixed)
    res = np.linalg.norm((F - ku)[freedofs]) / np.linalg.norm(F[freedofs])
    note("  oracle residual of the FP64 solution", res)
    assert res <= 1e-10
    assert c64 == pytest.approx(float(F @ u64), rel=1e-10)
    # the headline accuracy bar (c within 1e-6 of FP64), and u, dc at rtol 1e-8
    assert ec <= 1e-6
    assert eu <= 1e-6 and edc <= 1e-6
    pr.close()


def test_design_iterations_mixed_vs_fp64_256x128x128():
    note("design iterations 256x128x128")
    pr, F, solid, void, free = _headline_problem()
    x0 = np.full(pr.shape, 0.3)
    x0[solid], x0[void] = 1.0, 0.0
    target = 0.3 * int(free.sum())
    runs = {}
    for prec, rtol in (("fp64", 1e-10), ("mixed", 1e-8)):
        st = DesignState(pr, x0)
        pr.have_U = False
        pr.plan.filter_project(st.x, 1.0, 0.5, st.rf, st.rho, st.drho)
        kw = dict(precision=prec, rtol=rtol, maxit=20000, solver="mg")
        recs = [design_step(pr, st, 3.0, 1.0, 0.5, 0.2, target, kw) for _ in range(4)]
        runs[prec] = (recs, st.x.cpu().numpy())
        note("  ", prec, [(r["c"], r["solve_iters"], r["oc_steps"]) for r in recs])
    (rd, xd), (rm, xm) = runs["fp64"], runs["mixed"]
    # the bisection stops at a relative multiplier bracket of 1e-8, so the two
    # paths may keep neighbouring midpoints: volumes agree to ~1e-8
    for a, b in zip(rm, rd):
        assert a["c"] == pytest.approx(b["c"], rel=1e-6)
        assert a["vol"] == pytest.approx(b["vol"], abs=1e-7)
        assert a["change"] == pytest.approx(b["change"], abs=1e-6)
    assert np.abs(xm - xd).max() <= 1e-6
    pr.close()


def test_fifty_design_iterations_mixed_vs_fp64_256x128x128():
    """The headline grid's 50 full design iterations (beta 1 then 2, 25 each,
    the edge-load cantilever, volfrac 0.3, rmin 3) through top3d on the mixed
    path (FP32 multigrid inside FP64 flexible CG, rtol 1e-8) and the FP64 path
    (rtol 1e-10): compliance per iteration within 1e-6, designs within 1e-5 and
    the final compliances within 1e-6."""
    nelx, nely, nelz = 256, 128, 128
    note("50 design iterations 256x128x128")
    F = edge_load_force_vector(nelx, nely, nelz)
    kw = dict(volfrac=0.3, penal=3.0, rmin=3.0, disp_thres=0.5, force_vector=F, maxloop=25,
              tolx=1e-12, beta_schedule=(1.0, 2.0), save_history=True, history_frequency=25)
    from pytopo3d_amd import top3d
    runs = {}
    for prec, rtol in (("mixed", 1e-8), ("fp64", 1e-10)):
        cs = []
        rho, hist, c, _ = top3d(nelx, nely, nelz, precision=prec, solver_rtol=rtol,
                                iteration_callback=lambda d: cs.append(d["c"]), **kw)
        runs[prec] = (np.asarray(cs), rho, c, hist)
    (cm, rm, fm, hm), (cd, rd, fd, hd) = runs["mixed"], runs["fp64"]
    assert cm.size == cd.size == 50
    err = float(np.max(np.abs(cm - cd) / np.abs(cd)))
    note("  max rel c diff per iteration", err, "| final c", fm, fd,
         "| max design diff", float(np.abs(rm - rd).max()))
    assert err <= 1e-6
    assert np.abs(rm - rd).max() <= 1e-5
    assert fm == pytest.approx(fd, rel=1e-6)
    assert hm["iteration_history"] == hd["iteration_history"]
