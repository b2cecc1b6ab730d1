"""CPU restatement (numpy/scipy, FP64) of PyTopo3D's per-iteration SIMP path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle the HIP path is checked
against.  It is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` -- never by the ``pytopo3d_amd`` product.

Every function names the reference file:line whose behaviour it restates
(paths are relative to the PyTopo3D source tree).  The code is written from the
reference's documented behaviour, not copied: assembly is COO->CSR with
``scipy.sparse``, the linear solve is SciPy's SuperLU (``spsolve``), which is the
reference's own CPU fallback when PyPardiso is absent (pytopo3d/utils/solver.py:18-29).

How each part is pinned (see tests/test_oracle_*.py):
  * element_stiffness      closed forms (affine normal/shear energy, KE[0,0]),
                           reflection invariance, rigid-body null space
  * numbering / edof       corner order [1,2,4,3,5,6,8,7], brute-force node ids
  * filter_matrix          dense brute-force distances (rectangular grids)
  * heaviside / density    central differences, the 0.35 known answer,
                           clamp invariants
  * filter_chain_rule      central differences with a non-symmetric H
  * solve / energies       residual of the assembled system, c = F.U,
                           finite differences of the whole chain
  * oc_projected           volume target, protected values, scale invariance,
                           brute-force lambda scan
  * top3d                  golden trajectory written by tests/golden/make_golden.py
"""

from __future__ import annotations

import math
from typing import Any, Callable, Dict, MutableMapping, Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

__all__ = [
    "H8_CORNERS",
    "EMIN",
    "constitutive_matrix",
    "element_stiffness",
    "node_index",
    "element_dofs",
    "force_vector",
    "edge_load_force_vector",
    "support_dofs",
    "filter_matrix",
    "heaviside",
    "physical_density",
    "filter_chain_rule",
    "oc_projected",
    "element_energy",
    "stiffness_scale",
    "assemble_stiffness",
    "solve_displacement",
    "apply_stiffness_ebe",
    "diag_stiffness_ebe",
    "MultigridCG",
    "fixed_geometry_metrics",
    "fixed_geometry_compliance",
    "boundary_protection",
    "top3d",
]

# Local corner order of the H8 element as (x, y, z) offsets of the element's
# minimum corner (pytopo3d/utils/assembly.py:17-26).
H8_CORNERS = (
    (0, 0, 0),
    (0, 1, 0),
    (1, 1, 0),
    (1, 0, 0),
    (0, 0, 1),
    (0, 1, 1),
    (1, 1, 1),
    (1, 0, 1),
)

# SIMP void modulus; E0 = 1 lives inside KE (pytopo3d/core/optimizer.py:94,584-586).
EMIN = 1.0e-9


# --------------------------------------------------------------------------
# Element stiffness (pytopo3d/utils/stiffness.py)
# --------------------------------------------------------------------------
def constitutive_matrix(E_x, E_y=None, E_z=None, G_xy=0.4, G_yz=None, G_zx=None,
                        nu_xy=0.3, nu_yz=None, nu_zx=None) -> np.ndarray:
    """6x6 Voigt stiffness as the inverse of the orthotropic compliance.

    Follows pytopo3d/utils/stiffness.py:133-197: missing moduli and ratios take
    the x / xy values, the Voigt order is [xx, yy, zz, xy, yz, zx] and the
    normal-normal compliance couplings are -nu_xy/E_x, -nu_yz/E_y, -nu_zx/E_z.
    """
    E_y = E_x if E_y is None else E_y
    E_z = E_x if E_z is None else E_z
    G_yz = G_xy if G_yz is None else G_yz
    G_zx = G_xy if G_zx is None else G_zx
    nu_yz = nu_xy if nu_yz is None else nu_yz
    nu_zx = nu_xy if nu_zx is None else nu_zx
    compliance = np.diag([1.0 / E_x, 1.0 / E_y, 1.0 / E_z,
                          1.0 / G_xy, 1.0 / G_yz, 1.0 / G_zx])
    compliance[0, 1] = compliance[1, 0] = -nu_xy / E_x
    compliance[1, 2] = compliance[2, 1] = -nu_yz / E_y
    compliance[0, 2] = compliance[2, 0] = -nu_zx / E_z
    return np.linalg.inv(compliance)


def element_stiffness(material_params: Optional[Sequence[float]] = None,
                      elem_size: float = 1.0) -> np.ndarray:
    """24x24 H8 stiffness by 2x2x2 Gauss quadrature on a cube of side h.

    Follows pytopo3d/utils/stiffness.py:12-130 (``lk_H8``): natural-coordinate
    signs from the corner table, J = dN/dxi^T X, B rows [xx, yy, zz, xy, yz, zx],
    KE = sum_q B^T C B det(J) with unit weights.
    """
    params = () if material_params is None else tuple(material_params)
    C = constitutive_matrix(*params) if params else constitutive_matrix(1.0)
    signs = 2.0 * np.asarray(H8_CORNERS, dtype=float) - 1.0          # (8, 3)
    X = signs * (0.5 * elem_size)
    g = 1.0 / math.sqrt(3.0)
    KE = np.zeros((24, 24))
    for gz in (-g, g):
        for gy in (-g, g):
            for gx in (-g, g):
                pt = np.array([gx, gy, gz])
                lin = 1.0 + signs * pt                                   # (8, 3)
                dN = np.empty((8, 3))
                dN[:, 0] = 0.125 * signs[:, 0] * lin[:, 1] * lin[:, 2]
                dN[:, 1] = 0.125 * signs[:, 1] * lin[:, 0] * lin[:, 2]
                dN[:, 2] = 0.125 * signs[:, 2] * lin[:, 0] * lin[:, 1]
                J = dN.T @ X
                detJ = np.linalg.det(J)
                grad = np.linalg.solve(J.T, dN.T).T                     # (8, 3)
                B = np.zeros((6, 24))
                for a in range(8):
                    gx_, gy_, gz_ = grad[a]
                    c = 3 * a
                    B[0, c] = gx_
                    B[1, c + 1] = gy_
                    B[2, c + 2] = gz_
                    B[3, c] = gy_
                    B[3, c + 1] = gx_
                    B[4, c + 1] = gz_
                    B[4, c + 2] = gy_
                    B[5, c] = gz_
                    B[5, c + 2] = gx_
                KE += B.T @ C @ B * detJ
    return KE


# --------------------------------------------------------------------------
# Numbering, loads and supports (pytopo3d/utils/assembly.py)
# --------------------------------------------------------------------------
def node_index(iy, ix, iz, nelx, nely):
    """Fortran-order node id nid = iy + ix*(nely+1) + iz*(nelx+1)*(nely+1)
    (pytopo3d/utils/assembly.py:121-127)."""
    return iy + ix * (nely + 1) + iz * (nelx + 1) * (nely + 1)


def element_dofs(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """(nele, 24) 1-based global DOFs per element, corner order H8_CORNERS
    (pytopo3d/utils/assembly.py:351-401)."""
    e = np.arange(nelx * nely * nelz)
    iy = e % nely
    ix = (e // nely) % nelx
    iz = e // (nelx * nely)
    edof = np.empty((e.size, 24), dtype=np.int64)
    for a, (ox, oy, oz) in enumerate(H8_CORNERS):
        nid = node_index(iy + oy, ix + ox, iz + oz, nelx, nely)
        for c in range(3):
            edof[:, 3 * a + c] = 3 * nid + c + 1
    return edof


def force_vector(nelx, nely, nelz, force_field=None) -> np.ndarray:
    """Global load vector (pytopo3d/utils/assembly.py:78-184).

    Default: -1 on the z DOF of every node on the line x = nelx, z = 0 (all y).
    With a ``(nely, nelx, nelz, 3)`` element force field each element's force
    is split 1/8 onto its eight corners.
    """
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    F = np.zeros(ndof)
    if force_field is None:
        iy = np.arange(nely + 1)
        nid = node_index(iy, nelx, 0, nelx, nely)
        F[3 * nid + 2] = -1.0
        return F
    force_field = np.asarray(force_field, dtype=float)
    if force_field.shape != (nely, nelx, nelz, 3):
        raise ValueError(
            f"force_field has shape {force_field.shape}, but expected {(nely, nelx, nelz, 3)}")
    # Element-by-element in the reference's loop order (z, x, y outer to inner,
    # corners dz, dx, dy) so the floating-point accumulation order matches.
    for ez in range(nelz):
        for ex in range(nelx):
            for ey in range(nely):
                f = force_field[ey, ex, ez]
                if not np.any(f):
                    continue
                share = f / 8.0
                for dz in (0, 1):
                    for dx in (0, 1):
                        for dy in (0, 1):
                            nid = node_index(ey + dy, ex + dx, ez + dz, nelx, nely)
                            F[3 * nid] += share[0]
                            F[3 * nid + 1] += share[1]
                            F[3 * nid + 2] += share[2]
    return F


def edge_load_force_vector(nelx, nely, nelz, total=1.0) -> np.ndarray:
    """Cantilever edge load: total ``total`` in -y spread evenly over the nodes
    of the edge x = nelx, y = 0 (all z).  Not expressible through the
    reference's element force field; used for the large synthetic problems."""
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    F = np.zeros(ndof)
    iz = np.arange(nelz + 1)
    nid = node_index(0, nelx, iz, nelx, nely)
    F[3 * nid + 1] = -total / (nelz + 1)
    return F


def support_dofs(nelx, nely, nelz, support_mask=None) -> Tuple[np.ndarray, np.ndarray]:
    """(free, fixed) 0-based DOF arrays (pytopo3d/utils/assembly.py:224-348).

    Default (or an all-False mask): every DOF of the nodes on x = 0.  With a
    mask, every DOF of every corner of every masked element.
    """
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    fixed_node = np.zeros((nelz + 1, nelx + 1, nely + 1), dtype=bool)   # [z, x, y]
    use_default = support_mask is None
    if not use_default:
        support_mask = np.asarray(support_mask, dtype=bool)
        if support_mask.shape != (nely, nelx, nelz):
            raise ValueError(
                f"support_mask has shape {support_mask.shape}, but expected {(nely, nelx, nelz)}")
        use_default = not support_mask.any()
    if use_default:
        fixed_node[:, 0, :] = True
    else:
        m = np.transpose(support_mask, (2, 1, 0))                       # [z, x, y]
        for dz in (0, 1):
            for dx in (0, 1):
                for dy in (0, 1):
                    fixed_node[dz:dz + nelz, dx:dx + nelx, dy:dy + nely] |= m
    fixed_dof = np.repeat(fixed_node.ravel(), 3)     # node-major, C order == nid order
    allp = np.arange(ndof)
    return allp[~fixed_dof], allp[fixed_dof]


# --------------------------------------------------------------------------
# Density filter and projection (pytopo3d/utils/filter.py)
# --------------------------------------------------------------------------
def filter_matrix(nelx, nely, nelz, rmin) -> Tuple[sp.csr_matrix, np.ndarray]:
    """Conical filter H and row sums Hs (pytopo3d/utils/filter.py:122-228).

    Weight rmin - sqrt(di^2 + dj^2 + dk^2) over integer index offsets, kept
    when > 1e-9; truncated at the domain boundary; zero row sums become 1.
    """
    if min(nelx, nely, nelz) < 1:
        raise ValueError("nelx, nely, and nelz must all be positive")
    if not np.isfinite(rmin) or rmin <= 0.0:
        raise ValueError(f"rmin must be finite and positive, got {rmin}")
    nele = nelx * nely * nelz
    R = int(math.ceil(rmin))
    e = np.arange(nele)
    iy = e % nely
    ix = (e // nely) % nelx
    iz = e // (nelx * nely)
    rows, cols, vals = [], [], []
    for dk in range(-R, R + 1):
        for di in range(-R, R + 1):
            for dj in range(-R, R + 1):
                w = rmin - np.sqrt(float(di * di + dj * dj + dk * dk))
                if not w > 1e-9:
                    continue
                jy, jx, jz = iy + dj, ix + di, iz + dk
                ok = (jy >= 0) & (jy < nely) & (jx >= 0) & (jx < nelx) & (jz >= 0) & (jz < nelz)
                rows.append(e[ok])
                cols.append((jy + jx * nely + jz * nelx * nely)[ok])
                vals.append(np.full(int(ok.sum()), w))
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nele, nele)).tocsr()
    H.sort_indices()
    Hs = np.asarray(H.sum(axis=1)).ravel()
    Hs[Hs == 0] = 1.0
    return H, Hs


def heaviside(rho_filtered, beta, eta=0.5):
    """tanh projection and its unclipped derivative (pytopo3d/utils/filter.py:26-60)."""
    if not np.isfinite(beta) or beta <= 0:
        raise ValueError(f"beta must be positive, got {beta}")
    if not 0.0 < eta < 1.0:
        raise ValueError(f"eta must be between 0 and 1, got {eta}")
    t_eta = np.tanh(beta * eta)
    den = t_eta + np.tanh(beta * (1.0 - eta))
    t = np.tanh(beta * (rho_filtered - eta))
    rho = (t_eta + t) / den
    d = beta * (1.0 - t * t) / den
    return np.clip(rho, 0.0, 1.0), d


def physical_density(x, H, Hs, beta, eta, solid, void):
    """rho~ = (H x)/Hs, projection, protected clamp (pytopo3d/utils/filter.py:63-100).

    Arrays keep the design's (nely, nelx, nelz) shape; element vectors are the
    Fortran flattening.  Solid -> 1, void -> 0 (void last, so it wins) and the
    local derivative is zeroed on both.
    """
    shape = x.shape
    rf = (H @ np.asarray(x).ravel(order="F")) / Hs
    rf = rf.reshape(shape, order="F")
    rho, d = heaviside(rf, beta, eta)
    rho = np.where(solid, 1.0, rho)
    rho = np.where(void, 0.0, rho)
    d = np.where(solid | void, 0.0, d)
    return rf, rho, d


def filter_chain_rule(g, d, H, Hs):
    """H^T ((g * d)/Hs) in Fortran element order (pytopo3d/utils/filter.py:103-119)."""
    shape = g.shape
    w = (g * d).ravel(order="F") / Hs
    return (H.T @ w).reshape(shape, order="F")


# --------------------------------------------------------------------------
# Projected optimality criteria (pytopo3d/utils/oc_update.py:25-124)
# --------------------------------------------------------------------------
def oc_projected(x, dc_dx, dv_dx, move, target_free_volume, H, Hs, beta, eta,
                 free, solid, void, max_steps=80, tol=1e-8, trace=None):
    """Bisection on the multiplier of the projected free-volume constraint.

    The ratio -dc/max(dv, 1e-20) (floored at 0) is normalised by its maximum
    (all ones if that is 0); lambda starts in [0, 1e12]; each step builds the
    move-limited candidate, re-filters and re-projects it, sums rho over the
    free set and keeps the candidate closest to the target (strict <).  V >
    target raises the lower bound.  Stops when (hi-lo)/max(hi+lo, tiny) < tol.
    ``trace`` (a list) receives (lambda, V) per step.
    """
    if not np.isfinite(move) or move <= 0:
        raise ValueError(f"move must be positive, got {move}")
    if max_steps < 1:
        raise ValueError("max_bisection_iterations must be at least 1")
    xf = x[free]
    dv = np.maximum(dv_dx[free], 1e-20)
    ratio = np.maximum(-dc_dx[free] / dv, 0.0)
    scale = float(np.max(ratio)) if ratio.size else 0.0
    ratio = ratio / scale if scale > 0.0 else np.ones_like(ratio)
    lo, hi = 0.0, 1.0e12
    best_x, best_rho, best_err = x.copy(), None, float("inf")
    tiny = np.finfo(float).tiny
    for _ in range(max_steps):
        lam = 0.5 * (lo + hi)
        r = np.maximum(ratio / (lam + 1e-30), 1e-30)
        c = np.maximum(xf - move, np.minimum(xf + move, xf * np.sqrt(r)))
        c = np.clip(c, 0.0, 1.0)
        cand = x.copy()
        cand[free] = c
        cand[solid] = 1.0
        cand[void] = 0.0
        _, rho, _ = physical_density(cand, H, Hs, beta, eta, solid, void)
        V = float(np.sum(rho[free]))
        if trace is not None:
            trace.append((lam, V))
        err = abs(V - target_free_volume)
        if err < best_err:
            best_x, best_rho, best_err = cand, rho, err
        if V > target_free_volume:
            lo = lam
        else:
            hi = lam
        if (hi - lo) / max(hi + lo, tiny) < tol:
            break
    return best_x, best_rho


# --------------------------------------------------------------------------
# Energies, assembly, solve (pytopo3d/core/compliance.py, core/optimizer.py)
# --------------------------------------------------------------------------
def element_energy(U, edof, KE):
    """ce[e] = Ue . KE . Ue with 1-based edof (pytopo3d/core/compliance.py:69-75)."""
    Ue = U[edof - 1]
    return np.einsum("ei,ij,ej->e", Ue, KE, Ue)


def stiffness_scale(rho, penal):
    """Emin + (1 - Emin) rho^p (pytopo3d/core/optimizer.py:584-586)."""
    return EMIN + (1.0 - EMIN) * rho ** penal


def assemble_stiffness(rho_flat_F, penal, KE, nelx, nely, nelz) -> sp.csr_matrix:
    """K = sum_e s_e KE scattered through edof (pytopo3d/core/optimizer.py:584-591)."""
    edof = element_dofs(nelx, nely, nelz) - 1
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    s = stiffness_scale(rho_flat_F, penal)
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (s[:, None] * KE.ravel()[None, :]).ravel()
    return sp.coo_matrix((vals, (rows, cols)), shape=(ndof, ndof)).tocsr()


def solve_displacement(rho, penal, KE, F, free, nelx, nely, nelz,
                       solver: Optional[Callable] = None) -> np.ndarray:
    """U with U[fixed] = 0 and Kff Uf = Ff (pytopo3d/core/optimizer.py:591-594).

    The default solver is SciPy's SuperLU, the reference's CPU path when
    PyPardiso is not installed (pytopo3d/utils/solver.py:18-29).
    """
    K = assemble_stiffness(np.asarray(rho).ravel(order="F"), penal, KE, nelx, nely, nelz)
    Kff = K[free, :][:, free]
    U = np.zeros(F.size)
    U[free] = (solver or spla.spsolve)(Kff.tocsc(), F[free])
    return U


def apply_stiffness_ebe(u, rho_flat_F, penal, KE, nelx, nely, nelz, fixed=None,
                        chunk=1 << 20) -> np.ndarray:
    """y = K u element by element, without assembling K.

    The gather U_e = u[edof] and the product with KE are those of
    pytopo3d/core/compliance.py:73-74; scaling by s_e and scatter-adding
    through the same edof is the product with the assembled matrix of
    pytopo3d/core/optimizer.py:584-591.  With ``fixed`` (0-based DOFs) the
    product is that of the reduced system as the matrix-free operator defines
    it: fixed entries of u are dropped and the fixed rows of y are u itself.
    Elements are processed in chunks to bound memory at large grids.
    """
    u = np.asarray(u, dtype=float)
    ndof = u.size
    uf = u
    if fixed is not None:
        uf = u.copy()
        uf[fixed] = 0.0
    s = stiffness_scale(np.asarray(rho_flat_F, dtype=float), penal)
    y = np.zeros(ndof)
    nele = nelx * nely * nelz
    edof_all = element_dofs(nelx, nely, nelz)
    for e0 in range(0, nele, chunk):
        e1 = min(nele, e0 + chunk)
        edof = edof_all[e0:e1] - 1
        ye = (uf[edof] @ KE) * s[e0:e1, None]          # KE symmetric: (KE U_e)^T = U_e^T KE
        y += np.bincount(edof.ravel(), weights=ye.ravel(), minlength=ndof)
    if fixed is not None:
        y[fixed] = u[fixed]
    return y


def diag_stiffness_ebe(rho_flat_F, penal, KE, nelx, nely, nelz) -> np.ndarray:
    """diag(K) element by element (the diagonal of the matrix of
    pytopo3d/core/optimizer.py:584-591)."""
    s = stiffness_scale(np.asarray(rho_flat_F, dtype=float), penal)
    edof = element_dofs(nelx, nely, nelz) - 1
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    return np.bincount(edof.ravel(), weights=(s[:, None] * np.diag(KE)[None, :]).ravel(),
                       minlength=ndof)


def _interp_1d(nf: int) -> sp.csr_matrix:
    """Linear interpolation from ceil((nf - 1) / 2) + 1 coarse nodes to nf fine
    nodes (even fine node 2i <- coarse i; odd fine node <- mean of neighbours)."""
    nc = (nf - 1 + 1) // 2 + 1
    rows, cols, vals = [], [], []
    for i in range(nf):
        if i % 2 == 0:
            rows.append(i); cols.append(i // 2); vals.append(1.0)
        else:
            rows += [i, i]; cols += [(i - 1) // 2, (i + 1) // 2]; vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))


class MultigridCG:
    """Conjugate gradients on Kff with a Galerkin geometric-multigrid V-cycle.

    A callable ``(Kff, bf) -> Uf`` for ``solve_displacement`` / ``top3d`` that
    replaces SuperLU where a direct factorisation takes minutes (3-D fill):
    the same system solved to a relative residual of ``rtol`` (default 1e-12)
    of the assembled matrix itself.  Coarse operators are Galerkin products
    P^T A P of trilinear node interpolation (so they see the exact element
    scales), smoothing is Chebyshev on D^-1 A over [lmax/30, 1.1 lmax] (lmax
    by power iteration), the coarsest level (<= coarse_dofs DOFs) is factorised.
    The previous solution is the next initial guess (``warm=True``), which
    does not change the result beyond the tolerance.  Test infrastructure:
    independent of the HIP solver (different coarse operators, smoother and
    Krylov variant).
    """

    def __init__(self, nelx, nely, nelz, free, rtol=1e-12, maxiter=20000, degree=3,
                 warm=True, coarse_dofs=3000):
        self.rtol, self.maxiter, self.degree, self.warm = rtol, maxiter, degree, warm
        self.free = np.asarray(free)
        self.last = None
        self.iterations = []
        # interpolation chain on the full DOF set, restricted to free DOFs
        dims = (nelx + 1, nely + 1, nelz + 1)
        ndof = 3 * dims[0] * dims[1] * dims[2]
        keep = np.zeros(ndof, bool)
        keep[self.free] = True
        self.P = []
        while keep.sum() > coarse_dofs and max(dims) > 2:
            Px, Py, Pz = (_interp_1d(n) for n in dims)
            Pn = sp.kron(Pz, sp.kron(Px, Py), format="csr")           # y fastest
            Pd = sp.kron(Pn, sp.identity(3), format="csr")            # component fastest
            Pf = Pd[keep]
            # coarse DOFs that touch a free fine DOF, minus those sitting on a
            # fixed fine DOF (the Dirichlet boundary is not represented twice)
            ckeep = np.asarray(abs(Pf).sum(axis=0)).ravel() > 0
            cx, cy, cz = (Pi.shape[1] for Pi in (Px, Py, Pz))
            cn = np.arange(cx * cy * cz)
            iy, ix, iz = cn % cy, (cn // cy) % cx, cn // (cx * cy)
            fy, fx, fz = 2 * iy, 2 * ix, 2 * iz
            inside = (fy < dims[1]) & (fx < dims[0]) & (fz < dims[2])
            fn = np.where(inside, fy + fx * dims[1] + fz * dims[0] * dims[1], 0)
            on_fixed = np.repeat(inside, 3) & ~keep[np.repeat(3 * fn, 3) + np.tile(np.arange(3), cn.size)]
            ckeep &= ~on_fixed
            self.P.append(Pf[:, ckeep].tocsr())
            keep = ckeep
            dims = tuple(Pi.shape[1] for Pi in (Px, Py, Pz))

    def _setup(self, A):
        self.A = [A.tocsr()]
        for P in self.P:
            self.A.append((P.T @ self.A[-1] @ P).tocsr())
        self.dinv = [1.0 / Ai.diagonal() for Ai in self.A]
        self.lmax = []
        rng = np.random.default_rng(0)
        for Ai, di in zip(self.A[:-1], self.dinv[:-1]):
            v = rng.standard_normal(Ai.shape[0])
            lam = 1.0
            for _ in range(20):
                w = di * (Ai @ v)
                lam = float(np.linalg.norm(w) / np.linalg.norm(v))
                v = w / np.linalg.norm(w)
            self.lmax.append(1.1 * lam)
        self.coarse = spla.splu(self.A[-1].tocsc())

    def _smooth(self, l, b, x):
        A, di = self.A[l], self.dinv[l]
        hi = self.lmax[l]
        lo = hi / 30.0
        theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
        sigma = theta / delta
        rho_old = 1.0 / sigma
        d = di * (b - A @ x) / theta
        x = x + d
        for _ in range(1, self.degree):
            rho_new = 1.0 / (2.0 * sigma - rho_old)
            d = rho_new * rho_old * d + 2.0 * rho_new / delta * di * (b - A @ x)
            x = x + d
            rho_old = rho_new
        return x

    def _vcycle(self, l, b):
        if l == len(self.A) - 1:
            return self.coarse.solve(b)
        x = self._smooth(l, b, np.zeros_like(b))
        r = b - self.A[l] @ x
        x = x + self.P[l] @ self._vcycle(l + 1, self.P[l].T @ r)
        return self._smooth(l, b, x)

    def __call__(self, Kff, bf):
        self._setup(Kff)
        A = self.A[0]
        x = self.last.copy() if (self.warm and self.last is not None) else np.zeros_like(bf)
        r = bf - A @ x
        nb = np.linalg.norm(bf)
        if nb == 0.0:
            return np.zeros_like(bf)
        z = self._vcycle(0, r)
        p = z.copy()
        rz = r @ z
        it = 0
        while np.linalg.norm(r) > self.rtol * nb:
            if it >= self.maxiter:
                raise RuntimeError(f"oracle multigrid CG: no convergence in {it} iterations")
            q = A @ p
            alpha = rz / (p @ q)
            x += alpha * p
            r -= alpha * q
            z = self._vcycle(0, r)
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
            it += 1
        self.iterations.append(it)
        self.last = x.copy()
        return x


def fixed_geometry_metrics(xPhys, penal, material_params=None, elem_size=0.01,
                           force_field=None, support_mask=None, obstacle_mask=None,
                           protected_zone_mask=None, force_vector_override=None,
                           solver: Optional[Callable] = None) -> Dict[str, Any]:
    """One FE solve on a fixed field (pytopo3d/core/optimizer.py:54-252);
    ``solver`` as for solve_displacement (default SuperLU)."""
    nely, nelx, nelz = xPhys.shape
    shape = (nely, nelx, nelz)
    obstacle_mask = np.zeros(shape, bool) if obstacle_mask is None else obstacle_mask
    protected_zone_mask = np.zeros(shape, bool) if protected_zone_mask is None else protected_zone_mask
    xe = np.array(xPhys, dtype=float, copy=True)
    xe[protected_zone_mask] = 1.0
    xe[obstacle_mask] = 0.0
    F = (force_vector(nelx, nely, nelz, force_field) if force_vector_override is None
         else np.asarray(force_vector_override, dtype=float))
    loaded = np.where(np.abs(F) > 0.0)[0]
    nodes = np.unique(loaded // 3)
    Fd = [float(np.sum(np.abs(F[3 * nodes + c]))) if nodes.size else 0.0 for c in range(3)]
    F_total = float(np.sum(np.abs(F[loaded]))) if loaded.size else 0.0
    free, _ = support_dofs(nelx, nely, nelz, support_mask)
    KE = element_stiffness(material_params, elem_size)
    U = solve_displacement(xe, penal, KE, F, free, nelx, nely, nelz, solver)
    ce = element_energy(U, element_dofs(nelx, nely, nelz), KE).reshape(shape, order="F")
    c = float(((EMIN + xe ** penal * (1.0 - EMIN)) * ce).sum())
    uavg = [float(np.mean(U[3 * nodes + a])) if nodes.size else None for a in range(3)]
    k = []
    for a in range(3):
        u = uavg[a]
        k.append(float(Fd[a] / abs(u)) if (u is not None and u != 0.0 and Fd[a] > 0.0) else None)
    dominant = max(range(3), key=lambda a: Fd[a])
    return {
        "compliance": c,
        "ux_avg_load_patch": uavg[0], "uy_avg_load_patch": uavg[1], "uz_avg_load_patch": uavg[2],
        "k_avg_x": k[0], "k_avg_y": k[1], "k_avg_z": k[2], "k_avg": k[dominant],
        "F_total": F_total, "F_total_x": Fd[0], "F_total_y": Fd[1], "F_total_z": Fd[2],
        "loaded_x_dof_count": int(nodes.size), "loaded_y_dof_count": int(nodes.size),
        "loaded_z_dof_count": int(nodes.size),
    }


def fixed_geometry_compliance(xPhys, penal, **kw) -> float:
    """Compliance only (pytopo3d/core/optimizer.py:255-279)."""
    return float(fixed_geometry_metrics(xPhys, penal, **kw)["compliance"])


# --------------------------------------------------------------------------
# The optimizer loop (pytopo3d/core/optimizer.py:283-839)
# --------------------------------------------------------------------------
def boundary_protection(nelx, nely, nelz, obstacle_mask, protected_zone_mask,
                        force_field, support_mask, force_vector_override=None):
    """Protected-solid / void masks with load and support strips added
    (pytopo3d/core/optimizer.py:343-405).  Returns (solid, void, free)."""
    shape = (nely, nelx, nelz)
    void = np.zeros(shape, bool) if obstacle_mask is None else np.asarray(obstacle_mask, bool).copy()
    solid = (np.zeros(shape, bool) if protected_zone_mask is None
             else np.asarray(protected_zone_mask, bool).copy())
    if void.shape != shape:
        raise ValueError(f"obstacle_mask has shape {void.shape}, expected {shape}")
    if solid.shape != shape:
        raise ValueError(f"protected_zone_mask has shape {solid.shape}, expected {shape}")
    if force_vector_override is not None:
        solid |= _elements_touching_loaded_nodes(force_vector_override, nelx, nely, nelz)
    elif force_field is None:
        solid[:, -1, 0] = True
    else:
        if force_field.shape != shape + (3,):
            raise ValueError(f"force_field has shape {force_field.shape}, expected {shape + (3,)}")
        solid |= np.any(force_field != 0.0, axis=-1)
    if support_mask is None:
        solid[:, 0, :] = True
    else:
        if support_mask.shape != shape:
            raise ValueError(f"support_mask has shape {support_mask.shape}, expected {shape}")
        if np.any(support_mask):
            solid |= np.asarray(support_mask, bool)
        else:
            solid[:, 0, :] = True
    solid[void & solid] = False
    free = ~(void | solid)
    return solid, void, free


def _elements_touching_loaded_nodes(F, nelx, nely, nelz):
    loaded = (np.abs(np.asarray(F).reshape(-1, 3)) > 0).any(axis=1)
    ln = loaded.reshape(nelz + 1, nelx + 1, nely + 1)                   # [z, x, y]
    touch = np.zeros((nelz, nelx, nely), bool)
    for dz in (0, 1):
        for dx in (0, 1):
            for dy in (0, 1):
                touch |= ln[dz:dz + nelz, dx:dx + nelx, dy:dy + nely]
    return np.transpose(touch, (2, 1, 0))


def top3d(nelx, nely, nelz, volfrac, penal, rmin, disp_thres, material_params=None,
          elem_size=0.01, obstacle_mask=None, force_field=None, support_mask=None,
          tolx=0.01, maxloop=2000, save_history=False, history_frequency=10,
          use_gpu=False, protected_zone_mask=None, beta_schedule=(1.0, 2.0, 4.0, 8.0),
          projection_eta=0.5, move=0.2, diagnostics_out: Optional[MutableMapping] = None,
          force_vector_override=None, solver=None, record: Optional[Callable] = None,
          skip_final_eval=False):
    """Oracle of ``top3d`` (pytopo3d/core/optimizer.py:283-839).

    Extra keywords (not in the reference): ``force_vector_override`` (a full
    load vector, elements touching loaded nodes become protected solid),
    ``solver`` (callable(Kff, bf) -> Uf, default SuperLU), ``record``
    (callback(dict) per design iteration) and ``skip_final_eval``.
    """
    if min(nelx, nely, nelz) < 1:
        raise ValueError("nelx, nely, and nelz must all be positive")
    if not 0.0 < volfrac <= 1.0:
        raise ValueError(f"volfrac must be in (0, 1], got {volfrac}")
    if not np.isfinite(penal) or penal <= 0:
        raise ValueError(f"penal must be positive, got {penal}")
    if not np.isfinite(maxloop) or maxloop < 1:
        raise ValueError(f"maxloop must be at least 1, got {maxloop}")
    if not np.isfinite(tolx) or tolx <= 0:
        raise ValueError(f"tolx must be positive, got {tolx}")
    if not np.isfinite(history_frequency) or history_frequency < 1:
        raise ValueError(f"history_frequency must be at least 1, got {history_frequency}")
    if not np.isfinite(move) or move <= 0:
        raise ValueError(f"move must be positive, got {move}")
    if not 0.0 < projection_eta < 1.0:
        raise ValueError(f"projection_eta must be between 0 and 1, got {projection_eta}")
    betas = tuple(float(b) for b in beta_schedule)
    if not betas:
        raise ValueError("beta_schedule must contain at least one value")
    if any(not np.isfinite(b) or b <= 0.0 for b in betas):
        raise ValueError(f"all beta_schedule values must be finite and positive, got {betas}")

    shape = (nely, nelx, nelz)
    solid, void, free = boundary_protection(nelx, nely, nelz, obstacle_mask, protected_zone_mask,
                                            force_field, support_mask, force_vector_override)
    design_nele = int(free.sum())
    if design_nele == 0:
        raise ValueError("optimization requires at least one free design element")
    target = float(volfrac * design_nele)

    F = (force_vector(nelx, nely, nelz, force_field) if force_vector_override is None
         else np.asarray(force_vector_override, dtype=float))
    freedofs, _ = support_dofs(nelx, nely, nelz, support_mask)
    KE = element_stiffness(material_params, elem_size)
    edof = element_dofs(nelx, nely, nelz)
    H, Hs = filter_matrix(nelx, nely, nelz, rmin)
    history = ({"density_history": [], "iteration_history": [], "compliance_history": [],
                "beta_history": []} if save_history else None)

    x = np.full(shape, float(volfrac))
    x[solid] = 1.0
    x[void] = 0.0

    def bounds(beta):
        lo_d = np.zeros(shape)
        lo_d[solid] = 1.0
        hi_d = np.ones(shape)
        hi_d[void] = 0.0
        _, rl, _ = physical_density(lo_d, H, Hs, beta, projection_eta, solid, void)
        _, rh, _ = physical_density(hi_d, H, Hs, beta, projection_eta, solid, void)
        return float(np.mean(rl[free])), float(np.mean(rh[free]))

    total = 0
    c = float("nan")
    stages = []
    for beta in betas:
        change = float("inf")
        it = 0
        rf, rho, d = physical_density(x, H, Hs, beta, projection_eta, solid, void)
        vmin, vmax = bounds(beta)
        feasible = vmin - 1e-10 <= volfrac <= vmax + 1e-10
        if volfrac < vmin - 1e-10:
            raise ValueError(
                "projected free-volume target is infeasible at "
                f"beta={beta:g}: target={volfrac:.6f}, minimum achievable={vmin:.6f}. "
                "Protected-solid density spreads into the free region through the density filter.")
        while change > tolx and it < maxloop:
            it += 1
            total += 1
            rho_fea = rho
            U = solve_displacement(rho, penal, KE, F, freedofs, nelx, nely, nelz, solver)
            ce = element_energy(U, edof, KE).reshape(shape, order="F")
            c = float(np.sum((EMIN + (1.0 - EMIN) * rho ** penal) * ce))
            dc_drho = -penal * (1.0 - EMIN) * rho ** (penal - 1.0) * ce
            dc_dx = filter_chain_rule(dc_drho, d, H, Hs)
            dv_dx = filter_chain_rule(free.astype(float), d, H, Hs)
            x_new, rho_new = oc_projected(x, dc_dx, dv_dx, move, target, H, Hs, beta,
                                          projection_eta, free, solid, void)
            change = float(np.max(np.abs(x_new[free] - x[free])))
            if record is not None:
                record({"beta": beta, "iter": it, "total": total, "x_in": x, "rho": rho_fea,
                        "U": U, "ce": ce, "c": c, "dc_dx": dc_dx, "dv_dx": dv_dx,
                        "x_new": x_new, "rho_new": rho_new, "change": change})
            x, rho = x_new, rho_new
            if save_history and ((total - 1) % history_frequency == 0 or change <= tolx
                                 or it == maxloop):
                history["density_history"].append(np.asarray(rho_fea).copy())
                history["iteration_history"].append(total - 1)
                history["compliance_history"].append(c)
                history["beta_history"].append(beta)
            rf, rho, d = physical_density(x, H, Hs, beta, projection_eta, solid, void)
        vf = float(np.mean(rho[free]))
        gray = float(np.mean(((rho[free] > 0.05) & (rho[free] < 0.95)).astype(float)))
        stages.append({
            "beta": float(beta), "iterations": int(it), "converged": bool(change <= tolx),
            "physical_density_fraction": vf, "volume_fraction_error": vf - float(volfrac),
            "gray_fraction_005_095": gray, "final_change": float(change),
            "minimum_achievable_physical_fraction": vmin,
            "maximum_achievable_physical_fraction": vmax,
            "target_within_achievable_range": feasible,
        })

    final_beta = betas[-1]
    rf, rho, _ = physical_density(x, H, Hs, final_beta, projection_eta, solid, void)
    if skip_final_eval:
        return rho, history, float("nan"), None
    kw = dict(material_params=material_params, elem_size=elem_size, force_field=force_field,
              support_mask=support_mask, obstacle_mask=void, protected_zone_mask=solid,
              force_vector_override=force_vector_override, solver=solver)
    final_c = fixed_geometry_compliance(rho, penal, **kw)
    if save_history:
        history["density_history"].append(rho.copy())
        history["iteration_history"].append(total)
        history["compliance_history"].append(float(final_c))
        history["beta_history"].append(float(final_beta))
    rb = (rho >= 0.5).astype(float)
    rb[solid] = 1.0
    rb[void] = 0.0
    bin_c = fixed_geometry_compliance(rb, penal, **kw)
    delta = (bin_c - final_c) / final_c if final_c != 0.0 else None
    fv = rho[free]
    metrics = {
        "projection_enabled": True,
        "projection_beta": float(final_beta),
        "projection_beta_schedule": [float(b) for b in betas],
        "projection_eta": float(projection_eta),
        "penal": float(penal),
        "design_variable_fraction": float(np.mean(x[free])),
        "filtered_density_fraction": float(np.mean(rf[free])),
        "physical_density_fraction": float(np.mean(fv)),
        "gray_fraction_005_095": float(np.mean(((fv > 0.05) & (fv < 0.95)).astype(float))),
        "projected_compliance": float(final_c),
        "binary_compliance": float(bin_c),
        "binary_compliance_delta": None if delta is None else float(delta),
        "binary_threshold": 0.5,
        "target_free_volume_fraction": float(volfrac),
        "physical_volume_fraction_error": float(np.mean(fv) - volfrac),
        "projection_total_iterations": int(total),
        "projection_stage_summaries": stages,
        "projection_converged": all(s["converged"] for s in stages),
        "projection_target_within_achievable_range": all(
            s["target_within_achievable_range"] for s in stages),
    }
    if diagnostics_out is not None:
        diagnostics_out.clear()
        diagnostics_out.update(metrics)
    return rho, history, float(final_c), None
