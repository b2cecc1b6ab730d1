"""Host-side problem set-up: element stiffness, loads, supports and masks.

These run once per ``top3d`` call on the host (as in the reference) and
produce the inputs the device kernels consume: KE (24x24), the load vector F
in the reference's DOF numbering, per-node fixed-DOF bits and per-element
protected-solid / void flags.
"""

from __future__ import annotations

import math
from typing import Optional, Sequence, Tuple

import numpy as np

# (x, y, z) offset of each H8 local corner (pytopo3d/utils/assembly.py:17-26).
H8_NODE_OFFSETS = (
    (0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0),
    (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1),
)

SOLID = 1
VOID = 2


def make_C_matrix(E_x: float, E_y=None, E_z=None, G_xy=0.4, G_yz=None, G_zx=None,
                  nu_xy=0.3, nu_yz=None, nu_zx=None, normalize: bool = False) -> np.ndarray:
    """Orthotropic 6x6 Voigt stiffness [xx, yy, zz, xy, yz, zx]
    (pytopo3d/utils/stiffness.py:133-197)."""
    E_y = E_x if E_y is None else E_y
    E_z = E_x if E_z is None else E_z
    G_yz = G_xy if G_yz is None else G_yz
    G_zx = G_xy if G_zx is None else G_zx
    nu_yz = nu_xy if nu_yz is None else nu_yz
    nu_zx = nu_xy if nu_zx is None else nu_zx
    S = np.zeros((6, 6))
    S[np.arange(3), np.arange(3)] = (1.0 / E_x, 1.0 / E_y, 1.0 / E_z)
    S[0, 1] = S[1, 0] = -nu_xy / E_x
    S[1, 2] = S[2, 1] = -nu_yz / E_y
    S[0, 2] = S[2, 0] = -nu_zx / E_z
    S[np.arange(3, 6), np.arange(3, 6)] = (1.0 / G_xy, 1.0 / G_yz, 1.0 / G_zx)
    C = np.linalg.inv(S)
    return C / E_x if (normalize and E_x != 0) else C


def lk_H8(E_x: float = 1, E_y=None, E_z=None, G_xy: float = 0.4, G_yz=None, G_zx=None,
          nu_xy: float = 0.3, nu_yz=None, nu_zx=None, elem_size: float = 1.0) -> np.ndarray:
    """24x24 H8 element stiffness, 2x2x2 Gauss quadrature on a cube of side
    ``elem_size`` (same signature and result as pytopo3d/utils/stiffness.py:12-130)."""
    C = make_C_matrix(E_x, E_y, E_z, G_xy, G_yz, G_zx, nu_xy, nu_yz, nu_zx)
    sgn = 2.0 * np.asarray(H8_NODE_OFFSETS, dtype=float) - 1.0            # (8 corners, 3)
    coords = sgn * (0.5 * elem_size)
    g = 1.0 / math.sqrt(3.0)
    # Gauss points in the reference's enumeration order (x fastest, then y, then z)
    pts = np.array([[sx * g, sy * g, sz * g] for sz in (-1, 1) for sy in (-1, 1)
                    for sx in (-1, 1)])
    KE = np.zeros((24, 24))
    for q in pts:
        one = 1.0 + sgn * q                                                # (8, 3)
        dN = np.stack([0.125 * sgn[:, 0] * one[:, 1] * one[:, 2],
                       0.125 * sgn[:, 1] * one[:, 0] * one[:, 2],
                       0.125 * sgn[:, 2] * one[:, 0] * one[:, 1]], axis=1)
        J = dN.T @ coords
        detJ = np.linalg.det(J)
        if detJ <= 0:
            raise ValueError("Negative or zero Jacobian determinant.")
        G = np.linalg.solve(J.T, dN.T).T                                   # (8, 3)
        B = np.zeros((6, 24))
        B[0, 0::3], B[1, 1::3], B[2, 2::3] = G[:, 0], G[:, 1], G[:, 2]
        B[3, 0::3], B[3, 1::3] = G[:, 1], G[:, 0]
        B[4, 1::3], B[4, 2::3] = G[:, 2], G[:, 1]
        B[5, 0::3], B[5, 2::3] = G[:, 2], G[:, 0]
        KE += B.T @ C @ B * detJ
    return KE


def element_stiffness(material_params: Optional[Sequence[float]], elem_size: float) -> np.ndarray:
    """KE as ``top3d`` builds it (pytopo3d/core/optimizer.py:429-432)."""
    if material_params is None:
        return lk_H8(elem_size=elem_size)
    return lk_H8(*material_params, elem_size=elem_size)


def n_dof(nelx: int, nely: int, nelz: int) -> int:
    return 3 * (nelx + 1) * (nely + 1) * (nelz + 1)


def _node_grid(nelx, nely, nelz):
    """Node-id array indexed [z, x, y] (Fortran node numbering)."""
    return np.arange((nelx + 1) * (nely + 1) * (nelz + 1)).reshape(nelz + 1, nelx + 1, nely + 1)


def build_force_vector(nelx: int, nely: int, nelz: int, force_field=None) -> np.ndarray:
    """Global load vector (pytopo3d/utils/assembly.py:78-184).

    Default: -1 on the z DOF of the nodes of the line x = nelx, z = 0.  An
    element force field ``(nely, nelx, nelz, 3)`` is split 1/8 onto the corners;
    corner shifts are accumulated in the order that gives every node its
    contributions in the reference's element order, so F is bitwise the same.
    """
    ndof = n_dof(nelx, nely, nelz)
    F = np.zeros(ndof)
    if force_field is None:
        nid = _node_grid(nelx, nely, nelz)[0, nelx, :]
        F[3 * nid + 2] = -1.0
        return F
    ff = np.asarray(force_field, dtype=float)
    if ff.shape != (nely, nelx, nelz, 3):
        raise ValueError(f"force_field has shape {ff.shape}, but expected {(nely, nelx, nelz, 3)}")
    share = np.transpose(ff, (2, 1, 0, 3)) / 8.0                            # [z, x, y, c]
    nz = np.any(ff != 0.0, axis=-1).transpose(2, 1, 0)
    share = np.where(nz[..., None], share, 0.0)
    Fn = np.zeros((nelz + 1, nelx + 1, nely + 1, 3))
    # a node receives from corner (dz, dx, dy) the element (iz-dz, ix-dx, iy-dy);
    # the reference visits elements z, x, y ascending, i.e. corners descending.
    for dz in (1, 0):
        for dx in (1, 0):
            for dy in (1, 0):
                Fn[dz:dz + nelz, dx:dx + nelx, dy:dy + nely] += share
    return Fn.reshape(-1)


def edge_load_force_vector(nelx: int, nely: int, nelz: int, total: float = 1.0) -> np.ndarray:
    """Cantilever edge load: total ``total`` in -y spread evenly over the nodes of
    the edge x = nelx, y = 0 (all z).  Pass it to ``top3d(force_vector=...)``."""
    F = np.zeros(n_dof(nelx, nely, nelz))
    nid = _node_grid(nelx, nely, nelz)[:, nelx, 0]
    F[3 * nid + 1] = -total / (nelz + 1)
    return F


def support_node_flags(nelx: int, nely: int, nelz: int, support_mask=None) -> np.ndarray:
    """Per-node fixed-DOF bits (7 = x, y, z fixed) in node order
    (pytopo3d/utils/assembly.py:224-348): the x = 0 face by default or for an
    all-False mask, else every corner of every masked element."""
    fixed = np.zeros((nelz + 1, nelx + 1, nely + 1), dtype=bool)
    default = support_mask is None
    if not default:
        sm = np.asarray(support_mask, dtype=bool)
        if sm.shape != (nely, nelx, nelz):
            raise ValueError(f"support_mask has shape {sm.shape}, but expected {(nely, nelx, nelz)}")
        default = not sm.any()
    if default:
        fixed[:, 0, :] = True
    else:
        m = sm.transpose(2, 1, 0)
        for dz in (0, 1):
            for dx in (0, 1):
                for dy in (0, 1):
                    fixed[dz:dz + nelz, dx:dx + nelx, dy:dy + nely] |= m
    return np.where(fixed.reshape(-1), 7, 0).astype(np.uint8)


def elements_touching_loaded_nodes(F: np.ndarray, nelx: int, nely: int, nelz: int) -> np.ndarray:
    """(nely, nelx, nelz) mask of elements with a loaded corner node."""
    loaded = (np.abs(np.asarray(F, dtype=float).reshape(-1, 3)) > 0).any(axis=1)
    ln = loaded.reshape(nelz + 1, nelx + 1, nely + 1)
    touch = np.zeros((nelz, nelx, nely), dtype=bool)
    for dz in (0, 1):
        for dx in (0, 1):
            for dy in (0, 1):
                touch |= ln[dz:dz + nelz, dx:dx + nelx, dy:dy + nely]
    return touch.transpose(2, 1, 0)


def protected_masks(nelx, nely, nelz, obstacle_mask=None, protected_zone_mask=None,
                    force_field=None, support_mask=None, force_vector=None,
                    logger=None) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(solid, void, free) element masks (pytopo3d/core/optimizer.py:343-405):
    load and support strips become protected solid, void wins on overlap.
    With an explicit ``force_vector`` the elements touching loaded nodes are
    protected (the analogue of the force-field rule)."""
    shape = (nely, nelx, nelz)
    void = (np.zeros(shape, bool) if obstacle_mask is None
            else np.asarray(obstacle_mask, dtype=bool).copy())
    solid = (np.zeros(shape, bool) if protected_zone_mask is None
             else np.asarray(protected_zone_mask, dtype=bool).copy())
    if void.shape != shape:
        raise ValueError(f"obstacle_mask has shape {void.shape}, expected {shape}")
    if solid.shape != shape:
        raise ValueError(f"protected_zone_mask has shape {solid.shape}, expected {shape}")
    if force_vector is not None:
        solid |= elements_touching_loaded_nodes(force_vector, nelx, nely, nelz)
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
            solid |= np.asarray(support_mask, dtype=bool)
        else:
            solid[:, 0, :] = True
    overlap = void & solid
    if np.any(overlap):
        if logger is not None:
            logger.warning(
                "Protected-solid and protected-void masks overlap at %d elements; "
                "protected void takes precedence.", int(np.sum(overlap)))
        solid[overlap] = False
    free = ~(void | solid)
    return solid, void, free


def element_flags(solid: np.ndarray, void: np.ndarray) -> np.ndarray:
    """uint8 flags in element (Fortran) order: 1 = protected solid, 2 = void."""
    f = np.zeros(solid.shape, dtype=np.uint8)
    f[solid] |= SOLID
    f[void] |= VOID
    return f.ravel(order="F")


def _check_box(nelx, nely, nelz, x1, x2, y1, y2, z1, z2):
    if not (0 <= x1 < x2 <= nelx and 0 <= y1 < y2 <= nely and 0 <= z1 < z2 <= nelz):
        raise ValueError("Specified index range is out of bounds")


def build_force_field(nelx: int, nely: int, nelz: int, x1: int, x2: int, y1: int, y2: int,
                      z1: int, z2: int, F_x: float, F_y: float, F_z: float) -> np.ndarray:
    """(nely, nelx, nelz, 3) element force field of a preset: the force vector
    divided evenly over the elements of the box [y1:y2, x1:x2, z1:z2]
    (pytopo3d/utils/assembly.py:28-74)."""
    _check_box(nelx, nely, nelz, x1, x2, y1, y2, z1, z2)
    ff = np.zeros((nely, nelx, nelz, 3), dtype=float)
    n = (x2 - x1) * (y2 - y1) * (z2 - z1)
    ff[y1:y2, x1:x2, z1:z2] = (F_x / n, F_y / n, F_z / n)
    return ff


def build_support_mask(nelx: int, nely: int, nelz: int, x1: int, x2: int, y1: int, y2: int,
                       z1: int, z2: int) -> np.ndarray:
    """(nely, nelx, nelz) support mask of a preset box
    (pytopo3d/utils/assembly.py:186-222)."""
    _check_box(nelx, nely, nelz, x1, x2, y1, y2, z1, z2)
    m = np.zeros((nely, nelx, nelz), dtype=bool)
    m[y1:y2, x1:x2, z1:z2] = True
    return m


def protected_zone_mask_from_ranges(nelx: int, nely: int, nelz: int,
                                    zones: Sequence[Sequence[int]]) -> np.ndarray:
    """Union of protected-zone boxes (x1, x2, y1, y2, z1, z2), each set as
    mask[y1:y2, x1:x2, z1:z2] (the custom-mask path of main.py:158-168)."""
    m = np.zeros((nely, nelx, nelz), dtype=bool)
    for x1, x2, y1, y2, z1, z2 in zones:
        m[y1:y2, x1:x2, z1:z2] = True
    return m
