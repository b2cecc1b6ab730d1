"""Shared helpers of the GPU parity tests (inputs built on the host, the
oracle evaluated on the CPU, the HIP path called through the C ABI)."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

import oracle
from oracle.simp import _interp_1d
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


def coarsen_scales(s, nelx, nely, nelz):
    """Element scales of the next multigrid level: a quarter of the sum of the
    (up to) 8 children; returns them and the level's element counts."""
    S = s.reshape((nely, nelx, nelz), order="F")
    cx, cy, cz = (nelx + 1) // 2, (nely + 1) // 2, (nelz + 1) // 2
    out = np.zeros((cy, cx, cz))
    for oy in range(2):
        for ox in range(2):
            for oz in range(2):
                sub = S[oy::2, ox::2, oz::2]
                out[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += sub
    return 0.25 * out.ravel(order="F"), (cx, cy, cz)


def assemble_level(s, nelx, nely, nelz, KE):
    """sum_e s_e KE scattered with the reference's element DOFs (scipy CSR)."""
    edof = oracle.element_dofs(nelx, nely, nelz) - 1
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (s[:, None] * KE.ravel()[None, :]).ravel()
    return sp.coo_matrix((vals, (rows, cols)), shape=(ndof, ndof)).tocsr()


def interp_level(dims_f):
    """Trilinear interpolation from the next coarser level to level dims_f."""
    nx, ny, nz = (d + 1 for d in dims_f)
    Pn = sp.kron(_interp_1d(nz), sp.kron(_interp_1d(nx), _interp_1d(ny)),
                 format="csr")
    return sp.kron(Pn, sp.identity(3), format="csr")


def coarse_fixed_bits(fixed, dims_f, dims_c):
    """Node bits of the coarser level: the fine node at twice the coarse index."""
    nxf, nyf, nzf = (d + 1 for d in dims_f)
    nxc, nyc, nzc = (d + 1 for d in dims_c)
    n = np.arange(nxc * nyc * nzc)
    iy, ix, iz = n % nyc, (n // nyc) % nxc, n // (nxc * nyc)
    fy, fx, fz = 2 * iy, 2 * ix, 2 * iz
    inside = (fy < nyf) & (fx < nxf) & (fz < nzf)
    fn = np.where(inside, fy + fx * nyf + fz * nxf * nyf, 0)
    return np.where(inside, fixed[fn], 0).astype(np.uint8)


def dof_mask(bits):
    """Per-DOF booleans from per-node fixed bits."""
    return np.stack([(bits >> c) & 1 for c in range(3)], axis=1).ravel().astype(bool)


def with_identity_rows(A, fixed):
    """A with the rows and columns of the fixed DOFs replaced by identity ones
    (the V-cycle's level operators mask inputs and keep fixed DOFs)."""
    A = A.tolil()
    idx = np.flatnonzero(fixed)
    A[idx, :] = 0.0
    A[:, idx] = 0.0
    A[idx, idx] = 1.0
    return A.tocsr()


MID_MIN_EL = 32768   # coarse levels of more elements take mid_degree / mid_cheb_lo


def cheb_coefs(lam, level, opt, nele):
    """Chebyshev degree and coefficients of one smoothing pass on a level of
    nele elements (the plan's options): first step d = c0 D^-1 r, then
    d = c1 d + c2 D^-1 r."""
    if level < 2:
        deg, lo = opt["smooth_degree"], opt["cheb_lo"]
    elif nele > MID_MIN_EL:
        deg, lo = opt["mid_degree"], opt["mid_cheb_lo"]
    else:
        deg, lo = opt["coarse_degree"], opt["coarse_cheb_lo"]
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


def level_stack(nelx, nely, nelz, rho, KE, first, nlev, emin=1e-9, penal=3.0):
    """Host matrices of multigrid levels first .. nlev - 1, rediscretised like
    the plan's hierarchy from the fine design rho: dicts of A (identity rows on
    fixed DOFs), the fixed-DOF mask, the level number and the interpolation
    from the next level."""
    s = emin + (1.0 - emin) * rho ** penal
    fixed = support_node_flags(nelx, nely, nelz)
    dims = (nelx, nely, nelz)
    out = []
    for level in range(nlev):
        if level >= first:
            out.append({"A": with_identity_rows(assemble_level(s, *dims, KE), dof_mask(fixed)),
                        "fixed": dof_mask(fixed), "level": level, "dims": dims,
                        "P": interp_level(dims) if level + 1 < nlev else None})
        if level + 1 < nlev:
            s, d2 = coarsen_scales(s, *dims)
            fixed = coarse_fixed_bits(fixed, dims, d2)
            dims = d2
    return out


def host_cycle(lv, k, b, lam, opt):
    """The V-cycle from entry k of level_stack's list down to its last entry
    (the coarsest grid, solved exactly) and back: Chebyshev pre-smoothing from
    x = 0 with D = diag(A), the residual restricted by P^T (zero on the coarse
    fixed DOFs), the next level's cycle, the correction on the free DOFs, the
    post-smoothing."""
    if k == len(lv) - 1:
        return np.linalg.solve(lv[k]["A"].toarray(), b)
    A, fl, Pm = lv[k]["A"], lv[k]["fixed"], lv[k]["P"]
    deg, c0, c1, c2 = cheb_coefs(lam, lv[k]["level"], opt, int(np.prod(lv[k]["dims"])))
    dinv = 1.0 / A.diagonal()
    x = c0 * dinv * b
    d = x.copy()
    for j in range(1, deg):
        d = c1[j] * d + c2[j] * dinv * (b - A @ x)
        x = x + d
    r = np.where(fl, 0.0, b - A @ x)
    bc = Pm.T @ r
    bc[lv[k + 1]["fixed"]] = 0.0
    x = x + np.where(fl, 0.0, Pm @ host_cycle(lv, k + 1, bc, lam, opt))
    d = c0 * dinv * (b - A @ x)
    x = x + d
    for j in range(1, deg):
        d = c1[j] * d + c2[j] * dinv * (b - A @ x)
        x = x + d
    return x
