"""End-to-end parity of the HIP ``top3d`` and the autograd compliance.

* the reference's own trajectories (golden fixtures written from the reference
  by tests/golden/make_golden.py): per-iteration compliance, density snapshots,
  final compliance and diagnostics;
* the oracle's top3d on the cantilever edge load (force_vector keyword);
* the reference's optimizer tests restated for the HIP path
  (tests/test_projected_optimizer.py:128-303 of the reference);
* autograd d c / d x against central differences and the oracle's dc/dx.
"""

import json
import os

import numpy as np
import pytest
import torch

import oracle
from golden_cases import METRIC_CASES, check_metrics
from gpu_helpers import dev, elem_vec, note, to_dev
from pytopo3d_amd import (
    SIMPProblem,
    edge_load_force_vector,
    evaluate_fixed_geometry_compliance,
    evaluate_fixed_geometry_metrics,
    top3d,
)
from pytopo3d_amd.autograd import compliance
from pytopo3d_amd.problem import element_stiffness, protected_masks, support_node_flags

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden():
    with open(os.path.join(GOLDEN, "top3d_golden.json")) as fh:
        index = json.load(fh)
    out = []
    for name, meta in sorted(index.items()):
        z = np.load(os.path.join(GOLDEN, f"top3d_{name}.npz"))
        kw = dict(meta["inputs"])
        for k in z.files:
            if k.startswith("in_"):
                kw[k[3:]] = z[k]
        for k in ("beta_schedule", "material_params"):
            if k in kw:
                kw[k] = tuple(kw[k])
        out.append((name, kw, z, meta))
    return out


GOLD = _golden()


@pytest.mark.parametrize("name,kw,z,meta", GOLD, ids=[g[0] for g in GOLD])
def test_top3d_reproduces_reference_trajectory(name, kw, z, meta):
    note("golden", name)
    diag = {}
    rho, hist, c, ff = top3d(diagnostics_out=diag, solver_rtol=1e-14, **kw)
    assert ff is None
    np.testing.assert_allclose(rho, z["rho"], rtol=0, atol=1e-8)
    np.testing.assert_allclose(hist["compliance_history"], z["compliance_history"], rtol=1e-9)
    np.testing.assert_allclose(np.stack(hist["density_history"]), z["density_history"], atol=1e-8)
    assert hist["iteration_history"] == list(z["iteration_history"])
    assert hist["beta_history"] == list(z["beta_history"])
    assert c == pytest.approx(meta["final_compliance"], rel=1e-9)
    for key, want in meta["diagnostics"].items():
        got = diag[key]
        if key == "projection_stage_summaries":
            for gs, ws in zip(got, want):
                for k2, v2 in ws.items():
                    if isinstance(v2, float):
                        assert gs[k2] == pytest.approx(v2, rel=1e-7, abs=1e-9), (key, k2)
                    else:
                        assert gs[k2] == v2, (key, k2)
        elif isinstance(want, float) and key.startswith("binary"):
            assert got == pytest.approx(want, rel=1e-5), key   # ill-conditioned 0/1 design
        elif isinstance(want, float):
            assert got == pytest.approx(want, rel=1e-7, abs=1e-9), key
        else:
            assert got == want, key


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_top3d_edge_load_matches_oracle(precision):
    nelx, nely, nelz = 12, 6, 4
    note("edge-load top3d", precision)
    F = edge_load_force_vector(nelx, nely, nelz)
    kw = dict(nelx=nelx, nely=nely, nelz=nelz, volfrac=0.3, penal=3.0, rmin=1.5, disp_thres=0.5,
              elem_size=1.0, tolx=1e-12, maxloop=3, beta_schedule=(1.0, 2.0),
              save_history=True, history_frequency=1)
    d_o, d_g = {}, {}
    rho_o, h_o, c_o, _ = oracle.top3d(force_vector_override=F, diagnostics_out=d_o, **kw)
    rho_g, h_g, c_g, _ = top3d(force_vector=F, diagnostics_out=d_g, precision=precision,
                               solver_rtol=1e-12, **kw)
    np.testing.assert_allclose(h_g["compliance_history"], h_o["compliance_history"], rtol=1e-9)
    np.testing.assert_allclose(rho_g, rho_o, atol=1e-8)
    assert c_g == pytest.approx(c_o, rel=1e-9)
    assert d_g["physical_density_fraction"] == pytest.approx(0.3, abs=1e-7)


def test_top3d_symmetric_tensile_update():
    nelx, nely, nelz = 4, 12, 2
    note("tensile symmetry")
    ff = np.zeros((nely, nelx, nelz, 3))
    ff[nely - 1:nely, 0:nelx, 0:nelz, 1] = 1.0 / (nelx * nelz)
    sm = np.zeros((nely, nelx, nelz), bool)
    sm[0:1, 0:nelx, 0:nelz] = True
    pz = np.zeros((nely, nelx, nelz), bool)
    pz[:3] = True
    pz[-3:] = True
    rho, _, _, _ = top3d(nelx, nely, nelz, 0.5, 3.0, 2.0, 0.5, force_field=ff, support_mask=sm,
                         protected_zone_mask=pz, tolx=1e-12, maxloop=1, beta_schedule=(1.0,))
    free = rho[3:-3]
    np.testing.assert_allclose(free, free[::-1], rtol=0, atol=1e-3)


def test_top3d_continuation_bookkeeping():
    note("continuation")
    shape = (2, 2, 2)
    sm = np.zeros(shape, bool); sm[:, 0, :] = True
    ff = np.zeros(shape + (3,)); ff[0, 1, 0, 0] = 1.0
    ob = np.zeros(shape, bool); ob[1, 1, 1] = True
    diag = {}
    rho, hist, c, _ = top3d(2, 2, 2, 0.5, 3.0, 1.0, 0.5, force_field=ff, support_mask=sm,
                            obstacle_mask=ob, tolx=1e-12, maxloop=1, save_history=True,
                            history_frequency=1, beta_schedule=(1.0, 2.0), diagnostics_out=diag)
    assert np.all(rho[sm | np.any(ff != 0, axis=-1)] == 1.0)
    assert rho[ob].item() == 0.0
    np.testing.assert_allclose(hist["density_history"][-1], rho)
    assert hist["iteration_history"] == [0, 1, 2]
    assert hist["beta_history"] == [1.0, 2.0, 2.0]
    assert hist["compliance_history"][-1] == pytest.approx(c)
    for dens, cs in zip(hist["density_history"], hist["compliance_history"]):
        rc = evaluate_fixed_geometry_compliance(xPhys=dens, penal=3.0, force_field=ff,
                                                support_mask=sm, obstacle_mask=ob)
        assert cs == pytest.approx(rc, rel=1e-9)
    assert diag["projection_beta_schedule"] == [1.0, 2.0]
    assert diag["projection_total_iterations"] == 2
    assert diag["projection_converged"] is False
    for st in diag["projection_stage_summaries"]:
        assert st["converged"] is False
        assert st["volume_fraction_error"] == pytest.approx(0.0, abs=1e-7)
    assert diag["physical_density_fraction"] == pytest.approx(0.5, abs=1e-7)
    assert diag["projected_compliance"] == pytest.approx(c)
    want = (diag["binary_compliance"] - diag["projected_compliance"]) / diag["projected_compliance"]
    assert diag["binary_compliance_delta"] == pytest.approx(want)


def test_top3d_default_strips_and_infeasible_target():
    note("default strips")
    diag = {}
    rho, _, _, _ = top3d(3, 2, 2, 0.5, 3.0, 1.5, 0.5, support_mask=np.zeros((2, 3, 2), bool),
                         maxloop=1, beta_schedule=(1.0,), diagnostics_out=diag)
    assert np.all(rho[:, 0, :] == 1.0) and np.all(rho[:, -1, 0] == 1.0)
    st = diag["projection_stage_summaries"][0]
    assert st["minimum_achievable_physical_fraction"] > 0.0
    assert st["target_within_achievable_range"] is True
    assert diag["physical_density_fraction"] == pytest.approx(0.5, abs=1e-7)
    with pytest.raises(ValueError, match="minimum achievable"):
        top3d(2, 2, 2, 0.3, 3.0, 1.5, 0.5, maxloop=1, beta_schedule=(1.0,))


def test_fixed_geometry_metrics_match_oracle():
    note("metrics")
    rng = np.random.default_rng(12)
    shape = (4, 6, 3)
    x = rng.uniform(0.05, 1.0, shape)
    ff = np.zeros(shape + (3,)); ff[0, 5, :, 1] = -0.5; ff[1, 5, :, 2] = 0.25
    got = evaluate_fixed_geometry_metrics(x, 3.0, elem_size=1.0, force_field=ff)
    want = oracle.fixed_geometry_metrics(x, 3.0, elem_size=1.0, force_field=ff)
    for k, v in want.items():
        if v is None:
            assert got[k] is None, k
        elif isinstance(v, float):
            assert got[k] == pytest.approx(v, rel=1e-9), k
        else:
            assert got[k] == v, k


def _problem(nelx, nely, nelz, rmin, solid, void, F):
    KE = element_stiffness(None, 1.0)
    fixed = support_node_flags(nelx, nely, nelz, None)
    return SIMPProblem(nelx, nely, nelz, rmin, KE, F, fixed, solid, void)


def test_autograd_compliance_gradient_central_difference_2x2x2():
    note("autograd 2x2x2")
    shape = (2, 2, 2)
    F = oracle.force_vector(2, 2, 2)
    z = np.zeros(shape, bool)
    pr = _problem(2, 2, 2, 1.5, z, z, F)
    x0 = np.linspace(0.35, 0.65, 8)
    x = torch.tensor(x0, dtype=torch.float64, device=dev(), requires_grad=True)
    kw = dict(penal=3.0, beta=4.0, eta=0.5, rtol=1e-14)
    c = compliance(x, pr, **kw)
    c.backward()
    g = x.grad.cpu().numpy()
    for i in (0, 3, 7):
        up, lo = x0.copy(), x0.copy()
        up[i] += 1e-6
        lo[i] -= 1e-6
        cu = compliance(torch.tensor(up, device=dev()), pr, **kw).item()
        cl = compliance(torch.tensor(lo, device=dev()), pr, **kw).item()
        assert g[i] == pytest.approx((cu - cl) / 2e-6, rel=2e-4, abs=1e-5)
    pr.close()


def test_autograd_gradient_with_protected_regions_matches_oracle_and_fd():
    """Small analogue of the 96x48x48 differentiation case: a protected void
    cylinder along z, a protected solid slab at x in [0, 2), beta = 8."""
    nelx, nely, nelz, rmin = 24, 12, 6, 2.5
    note("autograd protected", nelx, nely, nelz)
    shape = (nely, nelx, nelz)
    iy, ix = np.meshgrid(np.arange(nely), np.arange(nelx), indexing="ij")
    cyl = ((ix - 12) ** 2 + (iy - 6) ** 2 <= 4)[:, :, None].repeat(nelz, axis=2)
    slab = np.zeros(shape, bool); slab[:, 0:2, :] = True
    F = edge_load_force_vector(nelx, nely, nelz)
    solid, void, free = protected_masks(nelx, nely, nelz, cyl, slab, None, None, F)
    pr = _problem(nelx, nely, nelz, rmin, solid, void, F)
    rng = np.random.default_rng(3)
    x0 = rng.uniform(0.1, 0.9, shape)
    x0[solid], x0[void] = 1.0, 0.0
    xv = elem_vec(x0)
    x = torch.tensor(xv, device=dev(), requires_grad=True)
    kw = dict(penal=3.0, beta=8.0, eta=0.5, rtol=1e-14)
    c = compliance(x, pr, **kw)
    c.backward()
    g = x.grad.cpu().numpy()
    # oracle dc/dx on the same design (SuperLU)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    _, rho, d = oracle.physical_density(x0, H, Hs, 8.0, 0.5, solid, void)
    KE = oracle.element_stiffness(None, 1.0)
    freedofs, _ = oracle.support_dofs(nelx, nely, nelz)
    U = oracle.solve_displacement(rho, 3.0, KE, F, freedofs, nelx, nely, nelz)
    ce = oracle.element_energy(U, oracle.element_dofs(nelx, nely, nelz), KE).reshape(shape, order="F")
    c_o = float(np.sum(oracle.stiffness_scale(rho, 3.0) * ce))
    g_o = elem_vec(oracle.filter_chain_rule(-3.0 * (1 - 1e-9) * rho ** 2.0 * ce, d, H, Hs))
    assert c.item() == pytest.approx(c_o, rel=1e-10)
    assert np.abs(g - g_o).max() <= 1e-9 * np.abs(g_o).max()
    idx = np.random.default_rng(33).choice(np.flatnonzero(elem_vec(free)), 12, replace=False)
    for i in idx:
        up, lo = xv.copy(), xv.copy()
        up[i] += 1e-6
        lo[i] -= 1e-6
        cu = compliance(torch.tensor(up, device=dev()), pr, **kw).item()
        cl = compliance(torch.tensor(lo, device=dev()), pr, **kw).item()
        fd = (cu - cl) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())
    pr.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solver_non_convergence_raises_and_frees_the_plan(precision):
    """A Krylov cap far below what the solve needs raises the reference's
    RuntimeError("CG failed to converge ...") (pytopo3d/utils/solver.py:57-58),
    and top3d releases the plan's device memory on that path too (the mixed
    plan holds its ~2 GiB direction ring by then)."""
    nelx, nely, nelz = 96, 48, 48
    note("non-convergence", precision)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0, _ = torch.cuda.mem_get_info()
    with pytest.raises(RuntimeError, match="CG failed to converge"):
        top3d(nelx, nely, nelz, 0.3, 3.0, 2.0, 0.5, elem_size=1.0, maxloop=2,
              beta_schedule=(1.0,), precision=precision, solver_rtol=1e-12, solver_maxit=2)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free1, _ = torch.cuda.mem_get_info()
    note("  device memory not returned (MB)", (free0 - free1) / 1e6)
    assert free0 - free1 < 64e6


def test_solve_reuse_rejects_other_design_buffer_and_new_masks():
    """TK_SOLVE_REUSE is accepted only for the set-up's own rho buffer, penal
    and precision with the masks unchanged since."""
    nelx, nely, nelz = 10, 6, 4
    note("reuse checks")
    F = edge_load_force_vector(nelx, nely, nelz)
    z = np.zeros((nely, nelx, nelz), bool)
    pr = _problem(nelx, nely, nelz, 1.5, z, z, F)
    rng = np.random.default_rng(8)
    rho = to_dev(rng.uniform(0.1, 1.0, pr.plan.nele))
    other = rho.clone()
    U = torch.zeros_like(pr.F)
    pr.plan.prepare(rho, 3.0)
    pr.plan.solve(rho, 3.0, pr.F, U, reuse=True)
    with pytest.raises(ValueError, match="TK_SOLVE_REUSE"):
        pr.plan.solve(other, 3.0, pr.F, U, reuse=True)
    pr.plan.prepare(rho, 3.0)
    from pytopo3d_amd.problem import element_flags
    pr.plan.set_masks(to_dev(element_flags(z, z), torch.uint8),
                      to_dev(support_node_flags(nelx, nely, nelz, None), torch.uint8))
    with pytest.raises(ValueError, match="TK_SOLVE_REUSE"):
        pr.plan.solve(rho, 3.0, pr.F, U, reuse=True)
    with pytest.raises(ValueError, match="precision"):
        top3d(nelx, nely, nelz, 0.3, 3.0, 1.5, 0.5, precision="fp32")
    pr.close()


def test_history_snapshots_are_the_per_iteration_densities():
    """save_history copies the snapshot densities to the host asynchronously
    (pinned buffers on a side stream, waited for only before their device
    buffer is reused): a sparse snapshot schedule gives exactly the matching
    entries of the every-iteration history, and those equal the densities the
    iterations solved with as reported to the per-iteration callback."""
    nelx, nely, nelz = 20, 10, 6
    note("history snapshots")
    F = edge_load_force_vector(nelx, nely, nelz)
    kw = dict(elem_size=1.0, tolx=1e-12, maxloop=7, beta_schedule=(1.0, 4.0), force_vector=F,
              save_history=True)
    seen = []
    _, h1, c1, _ = top3d(nelx, nely, nelz, 0.3, 3.0, 2.0, 0.5, history_frequency=1,
                         iteration_callback=lambda d: seen.append(d["rho"].cpu().numpy()), **kw)
    _, h4, c4, _ = top3d(nelx, nely, nelz, 0.3, 3.0, 2.0, 0.5, history_frequency=4, **kw)
    assert h1["iteration_history"] == list(range(14)) + [14]
    assert h4["iteration_history"] == [0, 4, 6, 8, 12, 13, 14]
    for it, d in zip(h4["iteration_history"], h4["density_history"]):
        assert d.shape == (nely, nelx, nelz) and d.dtype == np.float64
        np.testing.assert_array_equal(d, h1["density_history"][it])
    for it in range(14):
        np.testing.assert_array_equal(elem_vec(h1["density_history"][it]), seen[it])
    assert c1 == c4


def test_runner_execute_optimization_reproduces_reference_and_saves_result(tmp_path):
    """execute_optimization with the reference runner's keyword set
    (pytopo3d/runners/experiment.py:95-213, forwarded to top3d at :173-195)
    reproduces the reference's golden trajectory, returns its 5-tuple, and
    save_result writes the (nely, nelx, nelz) float64 design as .npy
    (pytopo3d/utils/results_manager.py:86-104)."""
    from pytopo3d_amd.runner import execute_optimization, save_result
    note("runner")
    name, kw, z, meta = [g for g in GOLD if g[0] == "default_bc_6x4x3"][0]
    args = dict(kw)
    args["create_animation"] = args.pop("save_history")
    args["animation_frequency"] = args.pop("history_frequency")
    diag = {}
    xPhys, hist, c, failure, run_time = execute_optimization(
        logger=None, combined_obstacle_mask=None, use_gpu=True, protected_zone_mask=None,
        diagnostics_out=diag, solver_rtol=1e-14, **args)
    assert failure is None and run_time > 0.0
    np.testing.assert_allclose(xPhys, z["rho"], atol=1e-8)
    np.testing.assert_allclose(hist["compliance_history"], z["compliance_history"], rtol=1e-9)
    assert c == pytest.approx(meta["final_compliance"], rel=1e-9)
    path = save_result(xPhys, str(tmp_path / "exp"))
    back = np.load(path)
    assert path.endswith("optimized_design.npy")
    assert back.dtype == np.float64 and back.shape == z["rho"].shape
    np.testing.assert_array_equal(back, xPhys)


@pytest.mark.parametrize("name,x,kw,want", METRIC_CASES, ids=[c[0] for c in METRIC_CASES])
def test_fixed_geometry_metrics_match_reference_fixtures(name, x, kw, want):
    """evaluate_fixed_geometry_metrics against the dicts the reference's own
    function returned on the same designs and inputs
    (tests/golden/make_metrics_golden.py, pytopo3d/core/optimizer.py:54-252)."""
    note("metrics golden", name)
    check_metrics(evaluate_fixed_geometry_metrics(x, **kw), want)
