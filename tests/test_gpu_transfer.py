"""The V-cycle's grid transfers (tk_transfer_apply) against the host's
trilinear interpolation matrix built from 1-D factors (gpu_helpers.interp_level):

* restriction: coarse = P^T fine for a fine field that is zero on the fine
  fixed DOFs, zero on the coarse fixed DOFs;
* prolongation: fine += P coarse on the fine free DOFs only (the paired kernel,
  one thread per two fine nodes along y).

Grids with odd and even node counts along y (the pair kernel's last pair with
and without its odd node), the default x = 0 support and a random support
mask, both level transitions of each grid; FP64 and FP32.
"""

import numpy as np
import pytest
import torch

from gpu_helpers import (coarse_fixed_bits, dev, dof_mask, interp_level, make_plan, note,
                         to_dev)
from pytopo3d_amd.problem import support_node_flags

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nelx,nely,nelz,random_support", [(36, 18, 12, False),
                                                           (37, 19, 13, True),
                                                           (128, 64, 64, True)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restriction_and_prolongation_match_the_interpolation_matrix(nelx, nely, nelz,
                                                                     random_support, precision):
    note("transfers", nelx, nely, nelz, "random support" if random_support else "x = 0 face",
         precision)
    rng = np.random.default_rng(nelx + 7 * nely)
    support = rng.random((nely, nelx, nelz)) < 0.02 if random_support else None
    P, _ = make_plan(nelx, nely, nelz, 1.5, support_mask=support)
    levels = [tuple(d) for d in P.levels()]
    fixed = support_node_flags(nelx, nely, nelz, support)
    dt = torch.float64 if precision == "fp64" else torch.float32
    tol = 1e-14 if precision == "fp64" else 2e-6
    for level in range(min(2, len(levels) - 1)):
        dims_f, dims_c = levels[level], levels[level + 1]
        fixed_c = coarse_fixed_bits(fixed, dims_f, dims_c)
        Pm = interp_level(dims_f)
        mf, mc = dof_mask(fixed), dof_mask(fixed_c)
        assert Pm.shape == (mf.size, mc.size)
        # restriction of a field that is zero on the fine fixed DOFs
        rf = np.where(mf, 0.0, rng.standard_normal(mf.size))
        want = np.where(mc, 0.0, Pm.T @ rf)
        out = torch.full((mc.size,), np.nan, dtype=dt, device=dev())
        P.transfer_apply(level, 0, to_dev(rf, dt), out, precision=precision)
        got = out.double().cpu().numpy()
        err_r = float(np.abs(got - want).max() / np.abs(want).max())
        assert np.all(got[mc] == 0.0)
        # prolongation added on the free fine DOFs
        xc = rng.standard_normal(mc.size)
        xf = rng.standard_normal(mf.size)
        want = np.where(mf, xf, xf + Pm @ xc)
        out = to_dev(xf, dt)
        P.transfer_apply(level, 1, to_dev(xc, dt), out, precision=precision)
        got = out.double().cpu().numpy()
        err_p = float(np.abs(got - want).max() / np.abs(want).max())
        assert np.array_equal(got[mf], xf[mf].astype(np.float32 if dt == torch.float32 else
                                                      np.float64))
        note(f"  level {level} {dims_f} -> {dims_c}: restriction {err_r:.1e}, "
             f"prolongation {err_p:.1e}")
        assert err_r <= tol and err_p <= tol
        fixed = fixed_c
    with pytest.raises(ValueError):
        P.transfer_apply(len(levels) - 1, 0, out, out, precision=precision)
    P.close()
