"""The allocation-free entry point tk_energy_noalloc against tk_energy.

tk_energy_noalloc runs the same kernels as tk_energy on a plan (the sector-form
energy kernel and the fixed-order two-phase reduction of c) with its reduction
scratch in the library's static device array, so ce, dc/drho and c must be
bitwise those of tk_energy.  Checked on two calls in a row, the second on a
smaller grid (the static scratch keeps the first call's contents), through the
C ABI, with c both read on the host and written to a device output; a grid
over the scratch limit is refused with status 1.
"""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import dev, make_plan, note, to_dev
from pytopo3d_amd import _native

pytestmark = pytest.mark.gpu


def _case(nelx, nely, nelz, seed):
    rng = np.random.default_rng(seed)
    nele = nelx * nely * nelz
    rho = to_dev(rng.uniform(0.01, 1.0, nele))
    U = to_dev(rng.standard_normal(3 * (nelx + 1) * (nely + 1) * (nelz + 1)))
    return rho, U


def test_energy_noalloc_equals_plan_energy_on_consecutive_calls():
    KE = oracle.element_stiffness(None, 1.0)
    for (nelx, nely, nelz), seed in (((40, 24, 18), 1), ((12, 7, 5), 2)):
        note("energy_noalloc", nelx, nely, nelz)
        rho, U = _case(nelx, nely, nelz, seed)
        P, _ = make_plan(nelx, nely, nelz, 1.5)
        ce_p, dc_p = torch.empty_like(rho), torch.empty_like(rho)
        c_p = P.energy(U, rho, 3.0, ce=ce_p, dc_drho=dc_p)
        P.close()
        ce_n, dc_n = torch.full_like(rho, np.nan), torch.full_like(rho, np.nan)
        c_dev = torch.full((1,), np.nan, dtype=torch.float64, device=dev())
        c_n = _native.energy_noalloc(nelx, nely, nelz, KE, U, rho, 3.0, ce=ce_n, dc_drho=dc_n,
                                     c_dev=c_dev)
        assert torch.equal(ce_n, ce_p) and torch.equal(dc_n, dc_p)
        assert c_n == c_p and c_dev.item() == c_p
        # device output only, no host synchronisation inside the call
        c_dev2 = torch.zeros_like(c_dev)
        assert _native.energy_noalloc(nelx, nely, nelz, KE, U, rho, 3.0, c_dev=c_dev2,
                                      sync=False) is None
        assert c_dev2.item() == c_p


def test_energy_noalloc_refuses_grids_over_its_scratch():
    KE = oracle.element_stiffness(None, 1.0)
    rho, U = _case(2, 2, 2, 3)
    L = _native.lib()
    ke = np.ascontiguousarray(KE)
    rc = L.tk_energy_noalloc(1024, 1024, 17, ke.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(rho.data_ptr()), 3.0,
                             None, None, None, None, None)
    assert rc == 1
    assert b"at most 16777216 elements" in L.tk_last_error()
    with pytest.raises(ValueError, match="mirror symmetry"):
        R = np.random.default_rng(0).standard_normal((24, 24))
        _native.energy_noalloc(2, 2, 2, KE + 1e-3 * (R + R.T), U, rho, 3.0)
