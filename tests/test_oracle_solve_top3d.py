"""Pins of the oracle's assembly, solve, sensitivities and optimizer loop (CPU).

Restates tests/test_projected_optimizer.py of the reference (:83-125 FD of
the full projected compliance gradient, :184-249 continuation bookkeeping,
:252-288 default boundary protection, :291-303 the infeasible-target error)
and checks the oracle's top3d trajectory against the golden fixtures written
by tests/golden/make_golden.py from the reference itself.
"""

import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from golden_cases import METRIC_CASES, check_metrics

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden_cases():
    with open(os.path.join(GOLDEN, "top3d_golden.json")) as fh:
        index = json.load(fh)
    for name, meta in sorted(index.items()):
        z = np.load(os.path.join(GOLDEN, f"top3d_{name}.npz"))
        kw = dict(meta["inputs"])
        for k in z.files:
            if k.startswith("in_"):
                kw[k[3:]] = z[k]
        for k in ("beta_schedule", "material_params"):
            if k in kw:
                kw[k] = tuple(kw[k])
        yield name, kw, z, meta


GOLDEN_CASES = list(_golden_cases())


def test_force_vector_default_and_field():
    F = oracle.force_vector(3, 2, 2)
    nid = oracle.node_index(np.arange(3), 3, 0, 3, 2)
    assert np.count_nonzero(F) == 3 and np.all(F[3 * nid + 2] == -1.0)
    ff = np.zeros((2, 3, 2, 3)); ff[1, 2, 1] = (8.0, -16.0, 4.0)
    F = oracle.force_vector(3, 2, 2, ff)
    assert F.reshape(-1, 3).sum(axis=0) == pytest.approx([8.0, -16.0, 4.0])
    assert np.count_nonzero(F.reshape(-1, 3)[:, 0]) == 8


def test_edge_load_sums_to_unit_force():
    F = oracle.edge_load_force_vector(6, 3, 4).reshape(-1, 3)
    assert F[:, 1].sum() == pytest.approx(-1.0)
    assert np.count_nonzero(F) == 5
    assert F[:, 0].sum() == 0.0 and F[:, 2].sum() == 0.0


def test_supports_default_and_mask():
    free, fixed = oracle.support_dofs(3, 2, 2)
    assert fixed.size == 3 * 3 * 3
    sm = np.zeros((2, 3, 2), bool); sm[0, 1, 0] = True
    _, fixed = oracle.support_dofs(3, 2, 2, sm)
    assert fixed.size == 24
    _, fixed_empty = oracle.support_dofs(3, 2, 2, np.zeros((2, 3, 2), bool))
    assert fixed_empty.size == 27


def test_direct_solve_residual_and_work_identity():
    rng = np.random.default_rng(0)
    nelx, nely, nelz = 5, 3, 3
    KE = oracle.element_stiffness(None, 1.0)
    rho = rng.uniform(0.01, 1, (nely, nelx, nelz))
    F = oracle.edge_load_force_vector(nelx, nely, nelz)
    free, fixed = oracle.support_dofs(nelx, nely, nelz)
    U = oracle.solve_displacement(rho, 3.0, KE, F, free, nelx, nely, nelz)
    K = oracle.assemble_stiffness(rho.ravel(order="F"), 3.0, KE, nelx, nely, nelz)
    r = (K @ U - F)[free]
    assert np.linalg.norm(r) < 1e-12 * np.linalg.norm(F)
    assert np.all(U[fixed] == 0.0)
    ce = oracle.element_energy(U, oracle.element_dofs(nelx, nely, nelz), KE)
    c = np.sum(oracle.stiffness_scale(rho.ravel(order="F"), 3.0) * ce)
    assert c == pytest.approx(F @ U, rel=1e-12)
    # K is symmetric positive semi-definite with the 6 rigid modes when unsupported
    assert abs((K - K.T)).max() < 1e-12


def _objective_and_grad(x, H, Hs, beta, nelx, nely, nelz):
    shape = (nely, nelx, nelz)
    fixed = np.zeros(shape, bool)
    _, rho, d = oracle.physical_density(x, H, Hs, beta, 0.5, fixed, fixed)
    KE = oracle.element_stiffness(None, 1.0)
    F = oracle.force_vector(nelx, nely, nelz)
    free, _ = oracle.support_dofs(nelx, nely, nelz)
    U = oracle.solve_displacement(rho, 3.0, KE, F, free, nelx, nely, nelz)
    ce = oracle.element_energy(U, oracle.element_dofs(nelx, nely, nelz), KE).reshape(shape, order="F")
    c = float(np.sum(oracle.stiffness_scale(rho, 3.0) * ce))
    g = oracle.filter_chain_rule(-3.0 * (1 - 1e-9) * rho ** 2.0 * ce, d, H, Hs)
    return c, g


def test_full_chain_gradient_matches_finite_difference():
    nelx = nely = nelz = 2
    shape = (2, 2, 2)
    H, Hs = oracle.filter_matrix(2, 2, 2, 1.5)
    x = np.linspace(0.35, 0.65, 8).reshape(shape, order="F")
    _, g = _objective_and_grad(x, H, Hs, 4.0, nelx, nely, nelz)
    for i in (0, 3, 7):
        up = x.ravel(order="F").copy(); lo = up.copy()
        up[i] += 1e-6; lo[i] -= 1e-6
        cu, _ = _objective_and_grad(up.reshape(shape, order="F"), H, Hs, 4.0, nelx, nely, nelz)
        cl, _ = _objective_and_grad(lo.reshape(shape, order="F"), H, Hs, 4.0, nelx, nely, nelz)
        assert g.ravel(order="F")[i] == pytest.approx((cu - cl) / 2e-6, rel=2e-4, abs=1e-5)


@pytest.mark.parametrize("name,kw,z,meta", GOLDEN_CASES, ids=[c[0] for c in GOLDEN_CASES])
def test_oracle_top3d_reproduces_reference_trajectory(name, kw, z, meta):
    diag = {}
    rho, hist, c, _ = oracle.top3d(diagnostics_out=diag, **kw)
    np.testing.assert_allclose(rho, z["rho"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(hist["compliance_history"], z["compliance_history"], rtol=1e-11)
    np.testing.assert_allclose(np.stack(hist["density_history"]), z["density_history"],
                               rtol=0, atol=1e-12)
    assert hist["iteration_history"] == list(z["iteration_history"])
    assert hist["beta_history"] == list(z["beta_history"])
    assert c == pytest.approx(meta["final_compliance"], rel=1e-11)
    for key, want in meta["diagnostics"].items():
        got = diag[key]
        if key == "projection_stage_summaries":
            for gs, ws in zip(got, want):
                for k2, v2 in ws.items():
                    if isinstance(v2, float):
                        assert gs[k2] == pytest.approx(v2, rel=1e-9, abs=1e-12), (key, k2)
                    else:
                        assert gs[k2] == v2
        elif isinstance(want, float) and key.startswith("binary"):
            # binary designs contain 1e-9 void regions: their solve is ill-conditioned
            assert got == pytest.approx(want, rel=1e-5)
        elif isinstance(want, float):
            assert got == pytest.approx(want, rel=1e-9, abs=1e-12), key
        else:
            assert got == want, key


def test_continuation_bookkeeping():
    shape = (2, 2, 2)
    sm = np.zeros(shape, bool); sm[:, 0, :] = True
    ff = np.zeros(shape + (3,)); ff[0, 1, 0, 0] = 1.0
    ob = np.zeros(shape, bool); ob[1, 1, 1] = True
    diag = {}
    rho, hist, c, _ = oracle.top3d(2, 2, 2, 0.5, 3.0, 1.0, 0.5, force_field=ff, support_mask=sm,
                                   obstacle_mask=ob, tolx=1e-12, maxloop=1, save_history=True,
                                   history_frequency=1, beta_schedule=(1.0, 2.0),
                                   diagnostics_out=diag)
    assert hist["iteration_history"] == [0, 1, 2]
    assert hist["beta_history"] == [1.0, 2.0, 2.0]
    assert diag["projection_total_iterations"] == 2 and diag["projection_converged"] is False
    for st in diag["projection_stage_summaries"]:
        assert st["volume_fraction_error"] == pytest.approx(0.0, abs=1e-7)
    for dens, cs in zip(hist["density_history"], hist["compliance_history"]):
        rc = oracle.fixed_geometry_compliance(dens, 3.0, force_field=ff, support_mask=sm,
                                              obstacle_mask=ob)
        assert cs == pytest.approx(rc, rel=1e-9)


def test_default_boundary_strips_are_protected():
    diag = {}
    rho, _, _, _ = oracle.top3d(3, 2, 2, 0.5, 3.0, 1.5, 0.5, support_mask=np.zeros((2, 3, 2), bool),
                                maxloop=1, beta_schedule=(1.0,), diagnostics_out=diag)
    assert np.all(rho[:, 0, :] == 1.0) and np.all(rho[:, -1, 0] == 1.0)
    st = diag["projection_stage_summaries"][0]
    assert st["minimum_achievable_physical_fraction"] > 0.0
    assert st["target_within_achievable_range"] is True


def test_infeasible_target_raises():
    with pytest.raises(ValueError, match="minimum achievable"):
        oracle.top3d(2, 2, 2, 0.3, 3.0, 1.5, 0.5, maxloop=1, beta_schedule=(1.0,))


@pytest.mark.parametrize("nelx,nely,nelz", [(3, 2, 2), (7, 5, 4), (12, 6, 5)])
def test_element_by_element_operator_equals_assembled_matrix(nelx, nely, nelz):
    """The oracle's assembly-free K u (gather, KE product, scatter) is the
    product with the assembled matrix, on the free rows of a support-strip
    problem, with the fixed rows equal to u; its diagonal is diag(K)."""
    rng = np.random.default_rng(nelx)
    shape = (nely, nelx, nelz)
    sm = np.zeros(shape, bool)
    sm[rng.integers(0, nely), 0, :] = True
    KE = oracle.element_stiffness(None, 1.0)
    rho = rng.uniform(0.01, 1.0, shape).ravel(order="F")
    K = oracle.assemble_stiffness(rho, 3.0, KE, nelx, nely, nelz)
    free, fixed = oracle.support_dofs(nelx, nely, nelz, sm)
    u = rng.standard_normal(K.shape[0])
    uf = u.copy()
    uf[fixed] = 0.0
    want = K @ uf
    want[fixed] = u[fixed]
    got = oracle.apply_stiffness_ebe(u, rho, 3.0, KE, nelx, nely, nelz, fixed, chunk=7)
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    # a transposed KE product (wrong operand order) is distinguishable
    KEa = KE + np.triu(np.ones((24, 24)), 1) * 1e-3
    bad = oracle.apply_stiffness_ebe(u, rho, 3.0, KEa, nelx, nely, nelz, fixed)
    assert np.abs(bad - want).max() > 1e-6 * np.abs(want).max()
    np.testing.assert_allclose(oracle.diag_stiffness_ebe(rho, 3.0, KE, nelx, nely, nelz),
                               K.diagonal(), rtol=1e-14)


@pytest.mark.parametrize("levels_dofs", [3000, 400])
def test_multigrid_cg_matches_direct_solve(levels_dofs):
    """The oracle's multigrid CG solves the assembled system to its tolerance:
    the same displacements as SuperLU on a SIMP design with contrast 1e-9
    (near-binary filtered field) at two hierarchy depths, and the same answer
    warm-started."""
    nelx, nely, nelz = 24, 12, 8
    shape = (nely, nelx, nelz)
    rng = np.random.default_rng(4)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, 2.0)
    z = np.zeros(shape, bool)
    _, rho, _ = oracle.physical_density(rng.uniform(0.2, 1.0, shape), H, Hs, 8.0, 0.5, z, z)
    KE = oracle.element_stiffness(None, 1.0)
    F = oracle.edge_load_force_vector(nelx, nely, nelz)
    free, _ = oracle.support_dofs(nelx, nely, nelz)
    K = oracle.assemble_stiffness(rho.ravel(order="F"), 3.0, KE, nelx, nely, nelz)
    Kff = K[free, :][:, free]
    ud = spla.spsolve(Kff.tocsc(), F[free])
    mg = oracle.MultigridCG(nelx, nely, nelz, free, coarse_dofs=levels_dofs)
    u = mg(Kff, F[free])
    assert len(mg.A) >= 2
    assert np.abs(u - ud).max() <= 1e-9 * np.abs(ud).max()
    assert np.linalg.norm(Kff @ u - F[free]) <= 1e-12 * np.linalg.norm(F[free])
    u2 = mg(Kff, F[free])        # warm start from the previous solution
    assert np.abs(u2 - ud).max() <= 1e-9 * np.abs(ud).max()
    assert mg.iterations[1] <= 2


@pytest.mark.parametrize("name,x,kw,want", METRIC_CASES, ids=[c[0] for c in METRIC_CASES])
def test_oracle_fixed_geometry_metrics_match_reference(name, x, kw, want):
    """The oracle's fixed_geometry_metrics against the dicts the reference's
    evaluate_fixed_geometry_metrics returned (tests/golden/make_metrics_golden.py)."""
    check_metrics(oracle.fixed_geometry_metrics(x, **kw), want)
