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

from gpu_helpers import host_cycle, level_stack, make_plan, note, to_dev

pytestmark = pytest.mark.gpu


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
    lv = level_stack(nelx, nely, nelz, rho, KE, level, len(levels))
    assert [d["dims"] for d in lv] == levels[level:]
    n = lv[0]["A"].shape[0]
    b = np.random.default_rng(5).standard_normal(n)
    b[lv[0]["fixed"]] = 0.0
    want = host_cycle(lv, 0, b, P.galerkin_info()["bound"], P.options)
    dt = torch.float64 if precision == "fp64" else torch.float32
    x = torch.full((n,), np.nan, dtype=dt, device=to_dev(b).device)
    P.tail_apply(to_dev(b, dt), x, precision=precision)
    got = x.double().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    note(f"  {n} DOFs, coarsest {lv[-1]['A'].shape[0]}: rel err {err:.2e}")
    assert err <= (1e-12 if precision == "fp64" else 1e-4)
    # fixed DOFs stay zero
    assert np.all(got[lv[0]["fixed"]] == 0.0)
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
