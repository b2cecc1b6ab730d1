"""The single-workgroup tail of the V-cycle (tk_tail.hip) against a host
restatement built from independently assembled level matrices.

The host cycle on the tail level l (second-coarsest) and the coarsest level c:
A_l = sum_e s_e KE with the level's scales (a quarter of the sum of the 8
children's, level by level from s = Emin + (1 - Emin) rho^3), identity rows
and masked columns at the level's fixed DOFs; Chebyshev smoothing of the
V-cycle's degree on [lo, 1] * lambda_bar from x = 0 (D = diag(A_l)), the
residual restricted with the transpose of the trilinear interpolation (zero on
the coarse fixed DOFs), the coarsest system solved exactly, the correction
added on the free DOFs, and the post-smoothing (degrees and intervals from the
plan's options).  Grids: 32x16x16 (tail level 2, the coarse degree, the
coarsest inverse in LDS for FP32), 16x8x8 (tail level 1,
fine degree 2) and 24x24x24 (216 tail elements and 1029 tail DOFs: two element
passes and two DOFs per thread; 192 coarsest DOFs, the inverse read from
global memory).
"""

import numpy as np
import pytest
import torch

from gpu_helpers import (assemble_level, coarse_fixed_bits, coarsen_scales, dof_mask, interp_level,
                         make_plan, note, to_dev)
from pytopo3d_amd.problem import support_node_flags

pytestmark = pytest.mark.gpu
EMIN = 1e-9


def _with_identity_rows(A, fixed):
    A = A.tolil()
    idx = np.flatnonzero(fixed)
    A[idx, :] = 0.0
    A[:, idx] = 0.0
    A[idx, idx] = 1.0
    return A.tocsr()


def _cheb_coefs(lam, level, opt):
    deg = opt["coarse_degree"] if level >= 2 else opt["smooth_degree"]
    lo = opt["coarse_cheb_lo"] if level >= 2 else opt["cheb_lo"]
    lmax, lmin = lam * opt["cheb_hi"], lam * lo
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho_old = 1.0 / sigma
    c1, c2 = [0.0], [0.0]
    for _ in range(1, deg):
        rho_new = 1.0 / (2.0 * sigma - rho_old)
        c1.append(rho_new * rho_old)
        c2.append(2.0 * rho_new / delta)
        rho_old = rho_new
    return deg, 1.0 / theta, c1, c2


def _host_tail(A, Ac, Pm, fl, fc, b, lam, level, opt):
    deg, c0, c1, c2 = _cheb_coefs(lam, level, opt)
    dinv = 1.0 / A.diagonal()
    x = c0 * dinv * b
    d = x.copy()
    for k in range(1, deg):
        d = c1[k] * d + c2[k] * dinv * (b - A @ x)
        x = x + d
    r = np.where(fl, 0.0, b - A @ x)
    bc = Pm.T @ r
    bc[fc] = 0.0
    xc = np.linalg.solve(Ac.toarray(), bc)
    x = x + np.where(fl, 0.0, Pm @ xc)
    d = c0 * dinv * (b - A @ x)
    x = x + d
    for k in range(1, deg):
        d = c1[k] * d + c2[k] * dinv * (b - A @ x)
        x = x + d
    return x


@pytest.mark.parametrize("nelx,nely,nelz,level", [(32, 16, 16, 2), (16, 8, 8, 1),
                                                  (24, 24, 24, 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tail_matches_host_cycle(nelx, nely, nelz, level, precision):
    note("tail", nelx, nely, nelz, precision)
    P, KE = make_plan(nelx, nely, nelz, 1.5)
    levels = [tuple(d) for d in P.levels()]
    note("  levels", levels, "tail level", P.tail_level())
    assert P.tail_level() == level == len(levels) - 2
    rho = np.random.default_rng(21).uniform(0.01, 1.0, nelx * nely * nelz)
    P.prepare(to_dev(rho), 3.0, precision=precision)
    s = EMIN + (1.0 - EMIN) * rho ** 3
    fixed = support_node_flags(nelx, nely, nelz)
    dims = (nelx, nely, nelz)
    for _ in range(level):
        s, d2 = coarsen_scales(s, *dims)
        fixed = coarse_fixed_bits(fixed, dims, d2)
        dims = d2
    assert dims == levels[level]
    sc, dc = coarsen_scales(s, *dims)
    fixc = coarse_fixed_bits(fixed, dims, dc)
    fl, fc = dof_mask(fixed), dof_mask(fixc)
    A = _with_identity_rows(assemble_level(s, *dims, KE), fl)
    Ac = _with_identity_rows(assemble_level(sc, *dc, KE), fc)
    Pm = interp_level(dims)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    b[fl] = 0.0
    want = _host_tail(A, Ac, Pm, fl, fc, b, P.galerkin_info()["bound"], level, P.options)
    dt = torch.float64 if precision == "fp64" else torch.float32
    x = torch.full((A.shape[0],), np.nan, dtype=dt, device=to_dev(b).device)
    P.tail_apply(to_dev(b, dt), x, precision=precision)
    got = x.double().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    note(f"  {A.shape[0]} DOFs, coarsest {Ac.shape[0]}: rel err {err:.2e}")
    assert err <= (1e-12 if precision == "fp64" else 1e-4)
    # fixed DOFs stay zero
    assert np.all(got[fl] == 0.0)
    P.close()


def test_plan_without_tail_reports_none():
    # 6x4x4: the second level (3x2x2) is already the coarsest, so no tail level
    P, _ = make_plan(6, 4, 4, 1.5)
    note("levels", P.levels(), "tail", P.tail_level())
    assert P.tail_level() == -1
    P.prepare(to_dev(np.full(6 * 4 * 4, 0.5)), 3.0)
    b = to_dev(np.zeros(3 * 7 * 5 * 5))
    with pytest.raises(Exception, match="tail"):
        P.tail_apply(b, torch.empty_like(b))
    P.close()
