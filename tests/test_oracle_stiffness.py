"""Pins of the oracle's element stiffness and numbering (CPU).

Restates tests/test_stiffness.py of the reference (corner order
:62-66, reflection invariance :69-78, affine normal and shear energies :81-116,
closed-form KE[0,0] :119-160) against the oracle, and adds the rigid-body null
space, which a dropped or mis-signed B term would break.
"""

from itertools import product

import numpy as np
import pytest

import oracle

ORTHO = (10.0, 20.0, 30.0, 4.0, 5.0, 6.0, 0.10, 0.12, 0.08)


def _coords(h=1.0):
    return np.asarray(oracle.H8_CORNERS, dtype=float) * h


def _reflection(axis):
    X = _coords()
    R = X.copy()
    R[:, axis] = 1.0 - R[:, axis]
    perm = [int(np.flatnonzero(np.all(X == r, axis=1))[0]) for r in R]
    T = np.zeros((24, 24))
    flip = np.eye(3)
    flip[axis, axis] = -1.0
    for a, b in enumerate(perm):
        T[3 * b:3 * b + 3, 3 * a:3 * a + 3] = flip
    return T


def test_corner_order_matches_reference_edof():
    edof = oracle.element_dofs(1, 1, 1)
    np.testing.assert_array_equal((edof[0, ::3] + 2) // 3, [1, 2, 4, 3, 5, 6, 8, 7])


def test_element_dofs_match_brute_force_node_ids():
    nelx, nely, nelz = 3, 2, 4
    edof = oracle.element_dofs(nelx, nely, nelz)
    for e in range(nelx * nely * nelz):
        iy, ix, iz = e % nely, (e // nely) % nelx, e // (nelx * nely)
        for a, (ox, oy, oz) in enumerate(oracle.H8_CORNERS):
            nid = (iy + oy) + (ix + ox) * (nely + 1) + (iz + oz) * (nelx + 1) * (nely + 1)
            assert edof[e, 3 * a] == 3 * nid + 1


@pytest.mark.parametrize("axis", range(3))
def test_reflection_invariance(axis):
    KE = oracle.element_stiffness(ORTHO)
    T = _reflection(axis)
    np.testing.assert_allclose(T.T @ KE @ T, KE, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("axis", range(3))
def test_affine_normal_strain_energy(axis):
    h = 0.7
    KE = oracle.element_stiffness(ORTHO, h)
    C = oracle.constitutive_matrix(*ORTHO)
    u = np.zeros((8, 3))
    u[:, axis] = _coords(h)[:, axis]
    assert u.ravel() @ KE @ u.ravel() == pytest.approx(h ** 3 * C[axis, axis], rel=1e-12)


@pytest.mark.parametrize("du,dx,G", [(0, 1, ORTHO[3]), (1, 2, ORTHO[4]), (2, 0, ORTHO[5])])
def test_affine_shear_energy(du, dx, G):
    h = 0.7
    KE = oracle.element_stiffness(ORTHO, h)
    u = np.zeros((8, 3))
    u[:, du] = _coords(h)[:, dx]
    assert u.ravel() @ KE @ u.ravel() == pytest.approx(h ** 3 * G, rel=1e-12)


def test_closed_form_diagonal_entry():
    h = 0.7
    KE = oracle.element_stiffness(ORTHO, h)
    C = oracle.constitutive_matrix(*ORTHO)
    s = 2.0 * np.asarray(oracle.H8_CORNERS[0], dtype=float) - 1.0
    g = 1.0 / np.sqrt(3.0)
    want = 0.0
    for xi, eta, zeta in product((-g, g), repeat=3):
        nat = 0.125 * np.array([s[0] * (1 + s[1] * eta) * (1 + s[2] * zeta),
                                s[1] * (1 + s[0] * xi) * (1 + s[2] * zeta),
                                s[2] * (1 + s[0] * xi) * (1 + s[1] * eta)])
        gr = nat * (2.0 / h)
        eps = np.array([gr[0], 0.0, 0.0, gr[1], 0.0, gr[2]])
        want += eps @ C @ eps * (h / 2.0) ** 3
    assert KE[0, 0] == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("params", [None, ORTHO])
def test_rigid_body_modes_are_in_the_null_space(params):
    KE = oracle.element_stiffness(params, 1.3)
    X = _coords(1.3)
    modes = []
    for c in range(3):
        t = np.zeros((8, 3)); t[:, c] = 1.0; modes.append(t.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):
        r = np.zeros((8, 3)); r[:, a] = -X[:, b]; r[:, b] = X[:, a]; modes.append(r.ravel())
    for m in modes:
        assert np.abs(KE @ m).max() < 1e-12 * np.abs(KE).max()
    ev = np.linalg.eigvalsh(KE)
    assert np.sum(ev > 1e-10 * ev.max()) == 18


def test_default_material_is_the_reference_default():
    C = oracle.constitutive_matrix(1.0)
    assert np.isfinite(C).all() and C.shape == (6, 6)
    assert C[3, 3] == pytest.approx(0.4)        # G_xy = 0.4, not E / 2(1 + nu)
    assert np.isfinite(oracle.element_stiffness(None, 0.01)).all()
