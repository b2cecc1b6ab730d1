"""The Galerkin operator of the third multigrid level (tk_galerkin.hip).

* Its product A_2 u against an independent one: the rediscretised
  second-level matrix A_1 = sum_e s1_e KE (s1 = a quarter of the sum of the
  8 children's fine scales, as the hierarchy coarsens) assembled with scipy,
  and A_2 = P^T A_1 P with the trilinear interpolation P, identity rows on the
  level's fixed DOFs (the fine fixed DOFs sampled at even nodes).  FP64 and
  FP32, on 64x32x32 and 72x40x36 (odd coarse counts), seeded uniform [0.01, 1]
  densities.  The second level's own product is checked the same way, and a
  plan created with the default options keeps the rediscretised third level.
* Solves with and without it agree to the tolerance, and on a design with
  sharp solid/void blocks it does not cost iterations (measured: 54 -> 46 in
  FP64, 49 -> 42 mixed).
"""

import numpy as np
import pytest
import torch

from gpu_helpers import (assemble_level as _assemble, coarse_fixed_bits as _coarse_fixed,
                         coarsen_scales as _coarsen, dev, dof_mask as _dof_mask,
                         interp_level as _interp, make_plan, note, to_dev)
from pytopo3d_amd import edge_load_force_vector
from pytopo3d_amd.problem import support_node_flags

pytestmark = pytest.mark.gpu
EMIN = 1e-9


def _check_level(P, level, A, fixed_bits, precision, seed, tol):
    mask = _dof_mask(fixed_bits)
    u = np.random.default_rng(seed).standard_normal(A.shape[0])
    um = np.where(mask, 0.0, u)
    want = A @ um
    want[mask] = u[mask]
    dt = torch.float64 if precision == "fp64" else torch.float32
    ud = to_dev(u, dt)
    yd = torch.empty_like(ud)
    P.level_apply(level, ud, yd, precision=precision)
    got = yd.double().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    note(f"  level {level} {precision} rel err {err:.2e}")
    assert err <= tol


@pytest.mark.parametrize("nelx,nely,nelz", [(64, 32, 32), (72, 40, 36)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_galerkin_level_operator_is_the_product(nelx, nely, nelz, precision):
    note("galerkin operator", nelx, nely, nelz, precision)
    P, KE = make_plan(nelx, nely, nelz, 1.5, options={"galerkin": 1})
    info = P.galerkin_info()
    levels = P.levels()
    note("  levels", levels, "galerkin", info)
    assert info["level"] == 2
    # the fixed bound max_c lambda_max(diag(M_c)^-1 M_c) (6.46 for this KE)
    assert 6.3 < info["bound"] < 6.6
    nele = nelx * nely * nelz
    rho = np.random.default_rng(11).uniform(0.01, 1.0, nele)
    P.prepare(to_dev(rho), 3.0, precision=precision)
    s0 = EMIN + (1.0 - EMIN) * rho ** 3
    s1, d1 = _coarsen(s0, nelx, nely, nelz)
    s2, d2 = _coarsen(s1, *d1)
    assert [tuple(levels[1]), tuple(levels[2])] == [d1, d2]
    f0 = support_node_flags(nelx, nely, nelz)
    f1 = _coarse_fixed(f0, (nelx, nely, nelz), d1)
    f2 = _coarse_fixed(f1, d1, d2)
    A1 = _assemble(s1, *d1, KE)
    A2 = (_interp(d1).T @ A1 @ _interp(d1)).tocsr()
    tol = 1e-13 if precision == "fp64" else 2e-6
    _check_level(P, 1, A1, f1, precision, 1, tol)
    _check_level(P, 2, A2, f2, precision, 2, tol)
    P.close()
    # by default (and with galerkin = 2) the third level is rediscretised like the others
    R, _ = make_plan(nelx, nely, nelz, 1.5, options={"galerkin": 2})
    assert R.galerkin_info()["level"] == -1
    R.close()
    Q, _ = make_plan(nelx, nely, nelz, 1.5)
    assert Q.galerkin_info()["level"] == -1
    Q.prepare(to_dev(rho), 3.0, precision=precision)
    _check_level(Q, 2, _assemble(s2, *d2, KE), f2, precision, 3, tol)
    Q.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_galerkin_level_solves_agree_and_save_iterations(precision):
    nelx, nely, nelz = 96, 48, 48
    note("galerkin solve", nelx, nely, nelz, precision)
    # voids and struts: a thresholded smooth random field (about 40 % void)
    rng = np.random.default_rng(4)
    field = rng.standard_normal((nely // 8 + 1, nelx // 8 + 1, nelz // 8 + 1))
    field = np.kron(field, np.ones((8, 8, 8)))[:nely, :nelx, :nelz]
    rho3 = np.where(field > -0.25, 1.0, 0.001)
    rho3[:, 0, :] = 1.0            # solid at the supports and along the loaded edge
    rho3[:, -2:, :2] = 1.0
    rho = rho3.ravel(order="F")
    Fh = edge_load_force_vector(nelx, nely, nelz)
    F = to_dev(Fh)
    rtol = 1e-10 if precision == "fp64" else 1e-8
    out = {}
    for g in (1, 0):
        P, _ = make_plan(nelx, nely, nelz, 1.5, options={"galerkin": g})
        U = torch.zeros(P.ndof, dtype=torch.float64, device=dev())
        it, rr = P.solve(to_dev(rho), 3.0, F, U, rtol=rtol, precision=precision)
        out[g] = (it, U.cpu().numpy())
        note(f"  compliance {float(Fh @ out[g][1]):.12e}")
        note(f"  galerkin option {g}: level {P.galerkin_info()['level']}, {it} iterations, "
             f"relres {rr:.2e}")
        P.close()
    (it_on, u_on), (it_off, u_off) = out[1], out[0]
    d = float(np.abs(u_on - u_off).max() / np.abs(u_off).max())
    dc = abs(float(Fh @ u_on) / float(Fh @ u_off) - 1.0)
    note(f"  displacement difference {d:.2e}, compliance difference {dc:.2e}")
    assert dc <= (1e-9 if precision == "fp64" else 1e-7)
    assert d <= (1e-6 if precision == "fp64" else 1e-4)
    assert it_on <= it_off
