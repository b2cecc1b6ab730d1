"""Shared helpers of the GPU parity tests (inputs built on the host, the
oracle evaluated on the CPU, the HIP path called through the C ABI)."""

from __future__ import annotations

import numpy as np
import torch

import oracle
from pytopo3d_amd import _native
from pytopo3d_amd.problem import element_flags, support_node_flags


def dev():
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev(), dtype=dtype)


def elem_vec(a3):
    """(nely, nelx, nelz) -> element-order vector."""
    return np.asarray(a3).ravel(order="F")


def make_plan(nelx, nely, nelz, rmin, solid=None, void=None, support_mask=None, KE=None,
              options=None):
    KE = oracle.element_stiffness(None, 1.0) if KE is None else KE
    P = _native.Plan(nelx, nely, nelz, KE, rmin, options)
    shape = (nely, nelx, nelz)
    solid = np.zeros(shape, bool) if solid is None else solid
    void = np.zeros(shape, bool) if void is None else void
    flags = to_dev(element_flags(solid & ~void, void), torch.uint8)
    fixed = to_dev(support_node_flags(nelx, nely, nelz, support_mask), torch.uint8)
    P.set_masks(flags, fixed)
    return P, KE


def random_masks(rng, shape, p_solid=0.1, p_void=0.1):
    solid = rng.random(shape) < p_solid
    void = rng.random(shape) < p_void
    return solid & ~void, void


def note(*args):
    print("[case]", *args, flush=True)
