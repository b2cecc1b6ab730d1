"""Device boundary-condition builders (tk_build_boundary_conditions) against
the oracle, bitwise: the load vector of build_force_vector
(pytopo3d/utils/assembly.py:78-184), the fixed DOFs of build_supports
(:224-348) and the protected-solid / void element masks of top3d
(pytopo3d/core/optimizer.py:343-405), for the default inputs, element force
fields with negative and mixed-sign entries, support masks (including an
all-False mask, which falls back to the x = 0 face), obstacles overlapping
protected solids, and the reference's preset boxes."""

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import dev, elem_vec, note
from pytopo3d_amd import _native
from pytopo3d_amd.problem import (build_force_field, build_support_mask, element_flags,
                                  protected_zone_mask_from_ranges)

pytestmark = pytest.mark.gpu


def _up(a, dtype):
    if a is None:
        return None
    a = np.asarray(a, dtype=bool) if dtype is np.uint8 else np.asarray(a)
    return torch.from_numpy(np.ascontiguousarray(a.astype(dtype))).to(dev())


def _cases():
    rng = np.random.default_rng(5)
    nelx, nely, nelz = 13, 9, 7
    shape = (nely, nelx, nelz)
    ff = np.zeros(shape + (3,))
    pick = rng.random(shape) < 0.15
    ff[pick] = rng.standard_normal((int(pick.sum()), 3))
    ff[pick & (rng.random(shape) < 0.3), 1] = 0.0       # some components exactly zero
    sm = rng.random(shape) < 0.05
    ob = rng.random(shape) < 0.1
    pz = rng.random(shape) < 0.1
    yield "default", (nelx, nely, nelz), None, None, None, None
    yield "field+mask", (nelx, nely, nelz), ff, sm, ob, pz
    yield "empty mask", (nelx, nely, nelz), ff, np.zeros(shape, bool), None, pz
    yield "obstacle only", (nelx, nely, nelz), None, sm, ob, None
    n2 = (40, 20, 10)
    yield "presets", n2, build_force_field(*n2, 38, 40, 0, 20, 4, 6, 0.0, -1.0, 0.5), \
        build_support_mask(*n2, 0, 2, 0, 20, 0, 10), None, \
        protected_zone_mask_from_ranges(*n2, [(10, 14, 0, 5, 0, 10), (20, 22, 15, 20, 3, 7)])


CASES = list(_cases())


@pytest.mark.parametrize("name,dims,ff,sm,ob,pz", CASES, ids=[c[0] for c in CASES])
def test_device_boundary_conditions_match_oracle_bitwise(name, dims, ff, sm, ob, pz):
    nelx, nely, nelz = dims
    note("device bcs", name, dims)
    P = _native.Plan(nelx, nely, nelz, oracle.element_stiffness(None, 1.0), 1.5)
    F = torch.full((P.ndof,), np.nan, dtype=torch.float64, device=dev())
    fixed = torch.full((P.nnode,), 255, dtype=torch.uint8, device=dev())
    flags = torch.full((P.nele,), 255, dtype=torch.uint8, device=dev())
    P.boundary_conditions(force_field=_up(ff, np.float64), support_mask=_up(sm, np.uint8),
                          obstacle_mask=_up(ob, np.uint8), protected_zone_mask=_up(pz, np.uint8),
                          F=F, node_fixed=fixed, elem_flags=flags)
    F_o = oracle.force_vector(nelx, nely, nelz, ff)
    assert np.array_equal(F.cpu().numpy(), F_o)
    assert F.cpu().numpy().tobytes() == F_o.tobytes()          # signed zeros too
    _, fixed_dofs = oracle.support_dofs(nelx, nely, nelz, sm)
    want = np.zeros(P.nnode, np.uint8)
    want[np.unique(fixed_dofs // 3)] = 7
    assert np.array_equal(fixed.cpu().numpy(), want)
    solid, void, _ = oracle.boundary_protection(nelx, nely, nelz, ob, pz, ff, sm)
    assert np.array_equal(flags.cpu().numpy(), element_flags(solid, void))
    P.close()
