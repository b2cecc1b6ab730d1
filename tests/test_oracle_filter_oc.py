"""Pins of the oracle's filter, projection, chain rule and OC update (CPU).

Restates tests/test_projection.py of the reference (:14-30 FD of the
projection derivative, :33-46 invalid parameters, :49-78 F-order identity and
clamps, :81-116 the 0.35 known answer, :119-184 the transposed chain rule with
a non-symmetric H, :187-210 the filter against brute-force distances,
:213-273 the projected OC volume target at three sensitivity scales) and adds a
brute-force lambda scan of the OC bisection.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import oracle


def test_projection_derivative_central_difference():
    rho = np.linspace(0.01, 0.99, 50)
    _, d = oracle.heaviside(rho, 8.0, 0.5)
    up, _ = oracle.heaviside(rho + 1e-7, 8.0, 0.5)
    lo, _ = oracle.heaviside(rho - 1e-7, 8.0, 0.5)
    np.testing.assert_allclose(d, (up - lo) / 2e-7, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("beta,eta", [(0.0, 0.5), (-1.0, 0.5), (np.nan, 0.5), (np.inf, 0.5),
                                      (1.0, 0.0), (1.0, 1.0)])
def test_projection_rejects_invalid_parameters(beta, eta):
    with pytest.raises(ValueError):
        oracle.heaviside(np.array([0.5]), beta, eta)


def test_projection_is_the_identity_at_the_threshold_limits():
    # rho(0) = 0, rho(1) = 1 and rho(eta) = tanh(b eta)/den for any beta
    for beta in (0.5, 1.0, 8.0, 64.0):
        r, _ = oracle.heaviside(np.array([0.0, 1.0]), beta, 0.3)
        np.testing.assert_allclose(r, [0.0, 1.0], atol=1e-15)


def test_physical_density_identity_filter_and_clamps():
    shape = (2, 2, 1)
    x = np.array([[[0.1], [0.3]], [[0.2], [0.4]]])
    H = sp.eye(4, format="csr")
    solid = np.zeros(shape, bool); void = np.zeros(shape, bool)
    solid[0, 1, 0] = True; void[1, 0, 0] = True
    rf, rho, d = oracle.physical_density(x, H, np.ones(4), 4.0, 0.5, solid, void)
    np.testing.assert_array_equal(rf.ravel(order="F"), x.ravel(order="F"))
    assert rho[solid].item() == 1.0 and rho[void].item() == 0.0
    assert d[solid].item() == 0.0 and d[void].item() == 0.0


def test_protected_values_filter_neighbours_known_answer():
    shape = (1, 3, 1)
    x = np.array([[[1.0], [0.2], [0.0]]])
    solid = np.zeros(shape, bool); void = np.zeros(shape, bool)
    solid[0, 0, 0] = True; void[0, 2, 0] = True
    H = sp.csr_matrix(np.array([[2.0, 1.0, 0.0], [1.0, 2.0, 1.0], [0.0, 1.0, 2.0]]))
    Hs = np.asarray(H.sum(axis=1)).ravel()
    rf, rho, d = oracle.physical_density(x, H, Hs, 4.0, 0.5, solid, void)
    assert rf[0, 1, 0] == pytest.approx(0.35)
    assert rho[solid].item() == 1.0 and rho[void].item() == 0.0
    assert d[0, 1, 0] > 0.0


def test_chain_rule_uses_the_transpose():
    shape = (2, 2, 1)
    x = np.array([[[0.25], [0.45]], [[0.35], [0.65]]])
    H = sp.csr_matrix(np.array([[2.0, 1.0, 0.0, 0.0], [0.2, 2.0, 0.8, 0.0],
                                [0.0, 0.4, 2.0, 0.6], [0.5, 0.0, 0.0, 2.0]]))
    Hs = np.asarray(H.sum(axis=1)).ravel()
    fixed = np.zeros(shape, bool)
    w = np.array([[[1.0], [0.4]], [[-0.2], [1.7]]])

    def obj(c):
        return float(np.sum(w * oracle.physical_density(c, H, Hs, 3.0, 0.5, fixed, fixed)[1]))

    _, _, d = oracle.physical_density(x, H, Hs, 3.0, 0.5, fixed, fixed)
    g = oracle.filter_chain_rule(w, d, H, Hs).ravel(order="F")
    fd = np.empty(4)
    for i in range(4):
        up = x.ravel(order="F").copy(); lo = up.copy()
        up[i] += 1e-7; lo[i] -= 1e-7
        fd[i] = (obj(up.reshape(shape, order="F")) - obj(lo.reshape(shape, order="F"))) / 2e-7
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)
    # the plain (non-transposed) product is distinguishable on this H
    wrong = (H @ ((w * d).ravel(order="F") / Hs))
    assert np.abs(wrong - fd).max() > 1e-3


@pytest.mark.parametrize("nelx,nely,nelz,rmin", [(3, 2, 4, 1.6), (4, 3, 2, 2.5), (5, 4, 3, 3.0),
                                                 (2, 2, 2, 1.0)])
def test_filter_matches_dense_brute_force(nelx, nely, nelz, rmin):
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    X = np.array([(ix, iy, iz) for iz in range(nelz) for ix in range(nelx) for iy in range(nely)],
                 dtype=float)
    D = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    want = np.maximum(0.0, rmin - D)
    want[want <= 1e-9] = 0.0
    np.testing.assert_allclose(H.toarray(), want)
    np.testing.assert_allclose(Hs, np.where(want.sum(axis=1) == 0, 1.0, want.sum(axis=1)))
    np.testing.assert_allclose(H.diagonal(), rmin)


@pytest.mark.parametrize("rmin,count", [(3.0, 93), (2.5, 81), (4.0, 251), (1.5, 19)])
def test_interior_stencil_sizes(rmin, count):
    H, _ = oracle.filter_matrix(11, 11, 11, rmin)
    centre = 5 + 5 * 11 + 5 * 121
    assert H.getrow(centre).nnz == count


@pytest.mark.parametrize("scale", (1.0, 1e-18, 1e18))
def test_projected_oc_hits_free_volume(scale):
    shape = (1, 6, 1)
    x = np.full(shape, 0.5)
    solid = np.zeros(shape, bool); void = np.zeros(shape, bool)
    solid[0, 0, 0] = True; void[0, -1, 0] = True
    free = ~(solid | void)
    x[solid] = 1.0; x[void] = 0.0
    H = sp.eye(6, format="csr"); Hs = np.ones(6)
    _, _, d = oracle.physical_density(x, H, Hs, 4.0, 0.5, solid, void)
    dc = oracle.filter_chain_rule(-scale * np.ones(shape), d, H, Hs)
    dv = oracle.filter_chain_rule(free.astype(float), d, H, Hs)
    xn, rn = oracle.oc_projected(x, dc, dv, 0.2, 0.4 * free.sum(), H, Hs, 4.0, 0.5, free,
                                 solid, void)
    assert xn[solid].item() == 1.0 and xn[void].item() == 0.0
    assert rn[solid].item() == 1.0 and rn[void].item() == 0.0
    assert np.mean(rn[free]) == pytest.approx(0.4, abs=1e-7)


def test_oc_bisection_agrees_with_a_brute_force_lambda_scan():
    rng = np.random.default_rng(4)
    nelx, nely, nelz, rmin, beta = 5, 4, 3, 1.5, 2.0
    shape = (nely, nelx, nelz)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    solid = np.zeros(shape, bool); void = np.zeros(shape, bool)
    free = ~(solid | void)
    x = rng.uniform(0.2, 0.8, shape)
    _, _, d = oracle.physical_density(x, H, Hs, beta, 0.5, solid, void)
    dc = oracle.filter_chain_rule(-rng.uniform(0.1, 1.0, shape), d, H, Hs)
    dv = oracle.filter_chain_rule(np.ones(shape), d, H, Hs)
    target = 0.45 * free.sum()
    trace = []
    xn, rn = oracle.oc_projected(x, dc, dv, 0.2, target, H, Hs, beta, 0.5, free, solid, void,
                                 trace=trace)
    # V(lambda) is non-increasing: a fine log-scan brackets the root the bisection found
    ratio = np.maximum(-dc / np.maximum(dv, 1e-20), 0.0)
    ratio /= ratio.max()

    def volume(lam):
        c = np.clip(np.maximum(x - 0.2, np.minimum(x + 0.2, x * np.sqrt(np.maximum(
            ratio / (lam + 1e-30), 1e-30)))), 0, 1)
        return np.sum(oracle.physical_density(c, H, Hs, beta, 0.5, solid, void)[1])

    lams = np.logspace(-6, 6, 2001)
    V = np.array([volume(l) for l in lams])
    assert np.all(np.diff(V) <= 1e-12)
    k = np.searchsorted(-V, -target)
    lam_best = min(trace, key=lambda t: abs(t[1] - target))[0]
    assert lams[k - 1] <= lam_best * (1 + 1e-6) and lam_best <= lams[k] * (1 + 1e-6)
    assert abs(np.sum(rn) - target) < 1e-6
    assert len(trace) <= 80
