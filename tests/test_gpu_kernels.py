"""Kernel-level parity of the HIP path against the CPU oracle (FP64 unless stated).

Bars: stencil neighbour sets and weights bit-exact; filtered / projected
densities within 1e-12; operator products within 1e-12 relative (FP32: 2e-6);
displacements, compliance and sensitivities within 1e-9 relative of the
oracle's SuperLU solve.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from gpu_helpers import dev, elem_vec, make_plan, note, random_masks, to_dev

pytestmark = pytest.mark.gpu

GRIDS = [(3, 2, 4, 1.6), (6, 4, 3, 1.5), (7, 5, 4, 2.5), (9, 6, 5, 3.0), (12, 5, 7, 2.0),
         (37, 66, 21, 2.4)]
# stencil reach 3 (rmin in (3, 4]): the tiled kernel's halo-3 instantiation;
# the second grid spans several tiles in every direction (tile 32 x 16 x 8)
HALO3 = [(9, 7, 6, 3.5), (20, 12, 10, 4.0), (70, 40, 19, 4.0)]


def _stencil_matrix(P, nelx, nely, nelz):
    off, w = P.stencil()
    nele = nelx * nely * nelz
    e = np.arange(nele)
    iy, ix, iz = e % nely, (e // nely) % nelx, e // (nelx * nely)
    rows, cols, vals = [], [], []
    for (dy, dx, dz), wt in zip(off, w):
        jy, jx, jz = iy + dy, ix + dx, iz + dz
        ok = (jy >= 0) & (jy < nely) & (jx >= 0) & (jx < nelx) & (jz >= 0) & (jz < nelz)
        rows.append(e[ok]); cols.append((jy + jx * nely + jz * nelx * nely)[ok])
        vals.append(np.full(ok.sum(), wt))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(nele, nele))


@pytest.mark.parametrize("nelx,nely,nelz,rmin", GRIDS + [(5, 5, 5, 4.0), (4, 3, 2, 1.0)])
def test_stencil_and_row_sums_match_oracle(nelx, nely, nelz, rmin):
    note("stencil", nelx, nely, nelz, rmin)
    P, _ = make_plan(nelx, nely, nelz, rmin)
    H = _stencil_matrix(P, nelx, nely, nelz)
    Ho, Hso = oracle.filter_matrix(nelx, nely, nelz, rmin)
    d = (H - Ho).tocoo()
    assert H.nnz == Ho.nnz
    assert d.nnz == 0 or np.max(np.abs(d.data)) == 0.0      # same neighbours, same weights
    Hs = torch.empty(P.nele, dtype=torch.float64, device=dev())
    P.row_sums(Hs)
    np.testing.assert_allclose(Hs.cpu().numpy(), Hso, rtol=1e-15, atol=0)
    P.close()


FILTER_KINDS = [0, 1]   # LDS tiles, stencil gather


@pytest.mark.parametrize("nelx,nely,nelz,rmin,halo,nsten", [(9, 6, 5, 3.0, 2, 93),
                                                           (9, 7, 6, 3.5, 3, 179),
                                                           (20, 12, 10, 4.0, 3, 251),
                                                           (6, 4, 3, 1.5, 1, 19),
                                                           (8, 8, 8, 5.0, 4, 485)])
def test_filter_kernel_variant_is_the_one_intended(nelx, nely, nelz, rmin, halo, nsten):
    """The plan takes the LDS-tiled kernel with the stencil's halo up to a
    reach of 3 (rmin <= 4) and the gather kernel beyond it or when asked, so
    the parity grids above exercise each instantiation they are meant to."""
    note("filter variant", rmin)
    for fk in FILTER_KINDS:
        P, _ = make_plan(nelx, nely, nelz, rmin, options={"filter_kind": fk})
        info = P.filter_info()
        assert info["halo"] == halo and info["nsten"] == nsten
        assert info["tiled"] == (fk == 0 and halo <= 3)
        P.close()


@pytest.mark.parametrize("fk", FILTER_KINDS)
@pytest.mark.parametrize("nelx,nely,nelz,rmin", GRIDS + HALO3)
@pytest.mark.parametrize("beta", [1.0, 8.0])
def test_filter_project_matches_oracle(nelx, nely, nelz, rmin, beta, fk):
    note("filter_project", nelx, nely, nelz, rmin, beta, "filter_kind", fk)
    rng = np.random.default_rng(nelx * 100 + nely)
    shape = (nely, nelx, nelz)
    solid, void = random_masks(rng, shape)
    P, _ = make_plan(nelx, nely, nelz, rmin, solid, void, options={"filter_kind": fk})
    x = rng.uniform(0.0, 1.0, shape)
    x[solid] = 1.0
    x[void] = 0.0
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    rf_o, rho_o, d_o = oracle.physical_density(x, H, Hs, beta, 0.5, solid, void)
    xd = to_dev(elem_vec(x))
    rf, rho, d = (torch.empty_like(xd) for _ in range(3))
    P.filter_project(xd, beta, 0.5, rf, rho, d)
    np.testing.assert_allclose(rf.cpu().numpy(), elem_vec(rf_o), rtol=0, atol=1e-14)
    np.testing.assert_allclose(rho.cpu().numpy(), elem_vec(rho_o), rtol=0, atol=1e-12)
    np.testing.assert_allclose(d.cpu().numpy(), elem_vec(d_o), rtol=1e-12, atol=1e-12)
    fixed = elem_vec(solid | void)
    assert np.all(d.cpu().numpy()[fixed] == 0.0)
    assert np.all(rho.cpu().numpy()[elem_vec(solid)] == 1.0)
    assert np.all(rho.cpu().numpy()[elem_vec(void)] == 0.0)
    P.close()


@pytest.mark.parametrize("fk", FILTER_KINDS)
@pytest.mark.parametrize("nelx,nely,nelz,rmin", GRIDS + HALO3)
def test_filter_transpose_matches_oracle_chain_rule(nelx, nely, nelz, rmin, fk):
    note("transpose2", nelx, nely, nelz, rmin, "filter_kind", fk)
    rng = np.random.default_rng(7 + nelz)
    shape = (nely, nelx, nelz)
    solid, void = random_masks(rng, shape)
    P, _ = make_plan(nelx, nely, nelz, rmin, solid, void, options={"filter_kind": fk})
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    g = rng.standard_normal(shape)
    dr = rng.uniform(0, 2, shape)
    dr[solid | void] = 0.0
    free = ~(solid | void)
    want_c = oracle.filter_chain_rule(g, dr, H, Hs)
    want_v = oracle.filter_chain_rule(free.astype(float), dr, H, Hs)
    gd, dd = to_dev(elem_vec(g)), to_dev(elem_vec(dr))
    oc, ov = torch.empty_like(gd), torch.empty_like(gd)
    P.filter_transpose2(gd, None, dd, oc, ov)
    np.testing.assert_allclose(oc.cpu().numpy(), elem_vec(want_c), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ov.cpu().numpy(), elem_vec(want_v), rtol=1e-12, atol=1e-14)
    P.close()


def _free_operator(rho, penal, KE, nelx, nely, nelz, support_mask=None):
    K = oracle.assemble_stiffness(elem_vec(rho), penal, KE, nelx, nely, nelz)
    free, fixed = oracle.support_dofs(nelx, nely, nelz, support_mask)
    return K, free, fixed


OP_GRIDS = [(3, 2, 2), (6, 4, 3), (9, 7, 5), (17, 16, 9), (31, 15, 12), (20, 33, 18)]


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("nelx,nely,nelz", OP_GRIDS)
def test_apply_K_matches_assembled_matrix(nelx, nely, nelz, kind):
    """FP64 and FP32 K u against the assembled matrix.  operator_kind 2 (BF16
    MFMA element contraction, FP32 accumulation) rounds u and KE to BF16 (8
    significant bits, unit roundoff 2^-8) in its FP32 sweeps: every entry is
    within 2^-6 of s |KE| |u| summed over its elements (twice the bound 2 * 2^-8
    of the two roundings; an indexing mistake is of order 1 of it); its FP64
    sweeps are those of kind 0."""
    note("apply_K", nelx, nely, nelz, "operator_kind", kind)
    rng = np.random.default_rng(3)
    shape = (nely, nelx, nelz)
    sm = np.zeros(shape, bool)
    sm[rng.integers(0, nely), 0, :] = True        # a support strip, not the default face
    P, KE = make_plan(nelx, nely, nelz, 1.5, support_mask=sm, options={"operator_kind": kind})
    rho = rng.uniform(0.01, 1.0, shape)
    K, free, fixed = _free_operator(rho, 3.0, KE, nelx, nely, nelz, sm)
    u = rng.standard_normal(3 * (nelx + 1) * (nely + 1) * (nelz + 1))
    want = np.empty_like(u)
    uf = u.copy()
    uf[fixed] = 0.0
    want[free] = (K @ uf)[free]
    want[fixed] = u[fixed]
    rd = to_dev(elem_vec(rho))
    ud = to_dev(u)
    y = torch.empty_like(ud)
    P.apply_K(rd, 3.0, ud, y)
    err = np.abs(y.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 1e-13, err
    y32 = torch.empty(u.size, dtype=torch.float32, device=dev())
    P.apply_K(rd, 3.0, ud.float(), y32, precision="mixed")
    if kind == 2:
        bound = oracle.apply_stiffness_ebe(np.abs(u), elem_vec(rho), 3.0, np.abs(KE), nelx, nely,
                                           nelz, fixed)
        diff = np.abs(y32.double().cpu().numpy() - want)[free]
        note("  bf16 error / bound", float((diff / bound[free]).max()))
        assert np.all(diff <= 2.0 ** -6 * bound[free])
    else:
        err32 = np.abs(y32.double().cpu().numpy() - want).max() / np.abs(want).max()
        assert err32 < 2e-6, err32
    diag = torch.empty_like(ud)
    P.diag_K(rd, 3.0, diag)
    dwant = K.diagonal().copy()
    dwant[fixed] = 1.0
    np.testing.assert_allclose(diag.cpu().numpy(), dwant, rtol=1e-14)
    P.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("solver", ["mg", "jacobi"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nelx,nely,nelz", [(6, 4, 3), (12, 6, 5), (16, 8, 8), (23, 17, 9)])
def test_solve_matches_direct_solve(nelx, nely, nelz, precision, solver, kind):
    note("solve", nelx, nely, nelz, precision, solver, "operator_kind", kind)
    rng = np.random.default_rng(11)
    shape = (nely, nelx, nelz)
    P, KE = make_plan(nelx, nely, nelz, 1.5, options={"operator_kind": kind})
    rho = rng.uniform(0.01, 1.0, shape)
    F = oracle.edge_load_force_vector(nelx, nely, nelz)
    free, _ = oracle.support_dofs(nelx, nely, nelz)
    U_o = oracle.solve_displacement(rho, 3.0, KE, F, free, nelx, nely, nelz)
    rd, Fd = to_dev(elem_vec(rho)), to_dev(F)
    U = torch.zeros_like(Fd)
    it, rr = P.solve(rd, 3.0, Fd, U, rtol=1e-13 if precision == "fp64" else 1e-12,
                     maxit=20000, precision=precision, jacobi=(solver == "jacobi"))
    note("  iters", it, "relres", rr)
    Ug = U.cpu().numpy()
    assert np.abs(Ug - U_o).max() / np.abs(U_o).max() < 1e-9
    c_o = float(F @ U_o)
    c = P.energy(U, rd, 3.0)
    assert abs(c - c_o) / c_o < 1e-10
    P.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_true_residual_across_tiles(precision):
    """The recursive FP64 residual of the Krylov loop must stay the true one:
    on a grid of 3 x 3 element-kernel tiles in x-y and several z chunks, the
    residual F - K U recomputed with the FP64 operator after the solve is at
    the requested tolerance.  (A direction formed with a different rounding in
    the owning and the neighbouring tile makes K p differ from the stored p;
    the true residual then stalls orders of magnitude above the recursive
    one.)"""
    nelx, nely, nelz = 40, 34, 20
    note("true residual", nelx, nely, nelz, precision)
    rng = np.random.default_rng(23)
    P, _ = make_plan(nelx, nely, nelz, 1.5)
    rd = to_dev(elem_vec(rng.uniform(0.01, 1.0, (nely, nelx, nelz))))
    Fd = to_dev(oracle.edge_load_force_vector(nelx, nely, nelz))
    U = torch.zeros_like(Fd)
    rtol = 1e-11
    it, rr = P.solve(rd, 3.0, Fd, U, rtol=rtol, maxit=20000, precision=precision)
    y = torch.empty_like(U)
    P.apply_K(rd, 3.0, U, y)
    true_rel = float(torch.linalg.vector_norm(Fd - y) / torch.linalg.vector_norm(Fd))
    note("  iters", it, "recursive", rr, "true", true_rel)
    assert rr <= rtol
    assert true_rel <= 10 * rtol, true_rel
    P.close()


@pytest.mark.parametrize("solver", ["mg", "jacobi"])
def test_mixed_direction_slots_give_the_same_solution(solver):
    """The mixed solver keeps the last `direction_slots` FP32 directions and
    adds them to the iterate every that many iterations and at the end: the
    same additions in the same order as one update per iteration, so U and the
    iteration count are bitwise the same for every slot count (folds during
    the solve at 2, 3 and 5 slots, only the final one at the default), cold
    and warm started; and U solves the system."""
    nelx, nely, nelz = 24, 12, 10
    note("direction slots", nelx, nely, nelz, solver)
    rng = np.random.default_rng(31)
    rd = to_dev(elem_vec(rng.uniform(0.01, 1.0, (nely, nelx, nelz))))
    Fd = to_dev(oracle.edge_load_force_vector(nelx, nely, nelz))
    jac = solver == "jacobi"
    outs = []
    for slots in (0, 2, 3, 5):
        P, _ = make_plan(nelx, nely, nelz, 1.5, options={"direction_slots": slots})
        U = torch.zeros_like(Fd)
        it, _ = P.solve(rd, 3.0, Fd, U, rtol=1e-9, maxit=20000, precision="mixed", jacobi=jac)
        U2 = 0.9 * U
        it2, _ = P.solve(rd, 3.0, Fd, U2, rtol=1e-11, maxit=20000, precision="mixed", jacobi=jac,
                         warm_start=True, reuse=True)
        note("  slots", slots, "iters", it, it2)
        outs.append((it, U.clone(), it2, U2.clone()))
        if slots == 0:
            y = torch.empty_like(U2)
            P.apply_K(rd, 3.0, U2, y)
            true_rel = float(torch.linalg.vector_norm(Fd - y) / torch.linalg.vector_norm(Fd))
            assert true_rel <= 1e-10, true_rel
        P.close()
    for it, U, it2, U2 in outs[1:]:
        assert it == outs[0][0] and it2 == outs[0][2]
        assert torch.equal(U, outs[0][1]) and torch.equal(U2, outs[0][3])
    for bad in (1, 65):
        with pytest.raises(ValueError, match="direction_slots"):
            make_plan(nelx, nely, nelz, 1.5, options={"direction_slots": bad})


@pytest.mark.parametrize("ke", ["h8", "asymmetric"])
@pytest.mark.parametrize("nelx,nely,nelz,rmin", [(6, 4, 3, 1.5), (10, 6, 4, 2.5)])
def test_energy_and_sensitivities_match_oracle(nelx, nely, nelz, rmin, ke):
    """The H8 KE takes the Walsh-Hadamard sector form of the energy; a KE
    without its mirror symmetry (a symmetric random perturbation) takes the
    dense 24x24 product."""
    note("energy", nelx, nely, nelz, rmin, ke)
    rng = np.random.default_rng(5)
    shape = (nely, nelx, nelz)
    solid, void = random_masks(rng, shape, 0.05, 0.05)
    KE = oracle.element_stiffness(None, 1.0)
    if ke == "asymmetric":
        R = rng.standard_normal((24, 24))
        KE = KE + 1e-3 * np.abs(KE).max() * (R + R.T)
    P, KE = make_plan(nelx, nely, nelz, rmin, solid, void, KE=KE)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    x = rng.uniform(0.05, 0.95, shape)
    x[solid], x[void] = 1.0, 0.0
    _, rho, d = oracle.physical_density(x, H, Hs, 4.0, 0.5, solid, void)
    U = rng.standard_normal(3 * (nelx + 1) * (nely + 1) * (nelz + 1))
    edof = oracle.element_dofs(nelx, nely, nelz)
    ce_o = oracle.element_energy(U, edof, KE).reshape(shape, order="F")
    c_o = float(np.sum(oracle.stiffness_scale(rho, 3.0) * ce_o))
    dc_o = -3.0 * (1 - 1e-9) * rho ** 2.0 * ce_o
    free = ~(solid | void)
    dcdx_o = oracle.filter_chain_rule(dc_o, d, H, Hs)
    dvdx_o = oracle.filter_chain_rule(free.astype(float), d, H, Hs)
    Ud, rd, dd = to_dev(U), to_dev(elem_vec(rho)), to_dev(elem_vec(d))
    ce, dc, dcdx, dvdx = (torch.empty_like(rd) for _ in range(4))
    c = P.energy(Ud, rd, 3.0, drho=dd, ce=ce, dc_drho=dc, dc_dx=dcdx, dv_dx=dvdx)
    np.testing.assert_allclose(ce.cpu().numpy(), elem_vec(ce_o), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dc.cpu().numpy(), elem_vec(dc_o), rtol=1e-12, atol=1e-12)
    assert abs(c - c_o) / abs(c_o) < 1e-12
    np.testing.assert_allclose(dcdx.cpu().numpy(), elem_vec(dcdx_o), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(dvdx.cpu().numpy(), elem_vec(dvdx_o), rtol=1e-12, atol=1e-14)
    P.close()


@pytest.mark.parametrize("fk", FILTER_KINDS)
@pytest.mark.parametrize("scale", [1.0, 1e-18, 1e18])
@pytest.mark.parametrize("nelx,nely,nelz,rmin,beta", [(6, 4, 3, 1.5, 1.0), (10, 6, 4, 2.5, 4.0),
                                                      (12, 8, 6, 3.0, 8.0), (40, 20, 35, 3.0, 2.0),
                                                      (9, 7, 6, 3.5, 2.0), (20, 12, 10, 4.0, 8.0)])
def test_oc_update_matches_oracle(nelx, nely, nelz, rmin, beta, scale, fk):
    note("oc", nelx, nely, nelz, rmin, beta, scale, "filter_kind", fk)
    rng = np.random.default_rng(17)
    shape = (nely, nelx, nelz)
    solid, void = random_masks(rng, shape, 0.05, 0.05)
    free = ~(solid | void)
    P, _ = make_plan(nelx, nely, nelz, rmin, solid, void, options={"filter_kind": fk})
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    x = rng.uniform(0.1, 0.9, shape)
    x[solid], x[void] = 1.0, 0.0
    _, rho, d = oracle.physical_density(x, H, Hs, beta, 0.5, solid, void)
    dc = -scale * rng.uniform(0.0, 1.0, shape)
    dc_dx = oracle.filter_chain_rule(dc, d, H, Hs)
    dv_dx = oracle.filter_chain_rule(free.astype(float), d, H, Hs)
    target = 0.4 * free.sum()
    trace = []
    xn_o, rn_o = oracle.oc_projected(x, dc_dx, dv_dx, 0.2, target, H, Hs, beta, 0.5, free,
                                     solid, void, trace=trace)
    xd = to_dev(elem_vec(x))
    xn = torch.empty_like(xd)
    steps = P.oc_update(xd, to_dev(elem_vec(dc_dx)), to_dev(elem_vec(dv_dx)), 0.2, target, beta,
                        0.5, xn, sync=True)
    assert steps == len(trace)
    np.testing.assert_allclose(xn.cpu().numpy(), elem_vec(xn_o), rtol=0, atol=1e-9)
    rho_n = torch.empty_like(xd)
    P.filter_project(xn, beta, 0.5, None, rho_n, None)
    P.design_stats(rho_n, xd, xn)
    res = P.results()
    assert abs(res["vol_sum"] - target) < 1e-6 * max(1.0, target)
    assert res["change"] == pytest.approx(np.max(np.abs(xn_o - x)[free]), abs=1e-9)
    P.close()


@pytest.mark.parametrize("beta", [1.0, 4.0])
def test_volume_bounds_match_oracle(beta):
    nelx, nely, nelz, rmin = 8, 5, 4, 2.0
    note("bounds", beta)
    rng = np.random.default_rng(2)
    shape = (nely, nelx, nelz)
    solid, void = random_masks(rng, shape, 0.1, 0.1)
    free = ~(solid | void)
    P, _ = make_plan(nelx, nely, nelz, rmin, solid, void)
    H, Hs = oracle.filter_matrix(nelx, nely, nelz, rmin)
    lo = np.zeros(shape); lo[solid] = 1.0
    hi = np.ones(shape); hi[void] = 0.0
    _, rl, _ = oracle.physical_density(lo, H, Hs, beta, 0.5, solid, void)
    _, rh, _ = oracle.physical_density(hi, H, Hs, beta, 0.5, solid, void)
    vmin, vmax = P.volume_bounds(beta, 0.5)
    assert vmin == pytest.approx(np.mean(rl[free]), abs=1e-13)
    assert vmax == pytest.approx(np.mean(rh[free]), abs=1e-13)
    P.close()


def test_reductions_are_bitwise_reproducible():
    nelx, nely, nelz = 40, 20, 16
    note("reproducible")
    rng = np.random.default_rng(9)
    P, _ = make_plan(nelx, nely, nelz, 2.0)
    rho = to_dev(rng.uniform(0.01, 1, nelx * nely * nelz))
    F = to_dev(oracle.edge_load_force_vector(nelx, nely, nelz))
    outs = []
    for _ in range(2):
        U = torch.zeros_like(F)
        P.solve(rho, 3.0, F, U, rtol=1e-10)
        outs.append((U.clone(), P.energy(U, rho, 3.0)))
    assert torch.equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    P.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_prepare_then_reuse_equals_solve(precision):
    """tk_prepare + a TK_SOLVE_REUSE solve gives bitwise the result of a plain
    solve, and a reuse after a set-up in another precision is refused."""
    nelx, nely, nelz = 14, 8, 6
    note("prepare/reuse", precision)
    rng = np.random.default_rng(21)
    P, _ = make_plan(nelx, nely, nelz, 1.5)
    rho = to_dev(rng.uniform(0.01, 1, nelx * nely * nelz))
    F = to_dev(oracle.edge_load_force_vector(nelx, nely, nelz))
    Ua, Ub = torch.zeros_like(F), torch.zeros_like(F)
    P.solve(rho, 3.0, F, Ua, rtol=1e-11, precision=precision)
    P.prepare(rho, 3.0, precision=precision)
    P.solve(rho, 3.0, F, Ub, rtol=1e-11, precision=precision, reuse=True)
    assert torch.equal(Ua, Ub)
    # a set-up builds the state of its own precision only
    other = "fp64" if precision == "mixed" else "mixed"
    P.prepare(rho, 3.0, precision=other)
    with pytest.raises(ValueError):
        P.solve(rho, 3.0, F, Ub, rtol=1e-11, precision=precision, reuse=True)
    with pytest.raises(ValueError):
        P.solve(rho, 2.0, F, Ub, rtol=1e-11, precision=other, reuse=True)
    P.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("solver", ["mg", "jacobi"])
def test_graph_replay_matches_direct_launches(precision, solver):
    """The captured Krylov chunks replay the same kernels with the same
    arguments: several solves in a row (different designs, warm starts, eager
    work and synchronisations in between) give bitwise the direct-launch
    results, and iteration counts agree."""
    nelx, nely, nelz = 18, 10, 8
    note("graph replay", precision, solver)
    rng = np.random.default_rng(5)
    F = to_dev(oracle.edge_load_force_vector(nelx, nely, nelz))
    plans = [make_plan(nelx, nely, nelz, 1.5, options={"no_graphs": g})[0] for g in (0, 1)]
    U = [torch.zeros_like(F), torch.zeros_like(F)]
    for k in range(4):
        rho = to_dev(rng.uniform(0.01, 1, nelx * nely * nelz))
        outs = []
        for P, u in zip(plans, U):
            it, rr = P.solve(rho, 3.0, F, u, rtol=1e-10, maxit=4000, precision=precision,
                             jacobi=(solver == "jacobi"), warm_start=(k % 2 == 1))
            torch.cuda.synchronize()
            outs.append((it, u.clone(), P.energy(u, rho, 3.0)))
        assert outs[0][0] == outs[1][0]
        assert torch.equal(outs[0][1], outs[1][1])
        assert outs[0][2] == outs[1][2]
    for P in plans:
        P.close()
