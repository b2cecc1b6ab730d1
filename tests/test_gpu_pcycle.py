"""The persistent coarse cycle (tk_pcycle.hip).

* Its part of the V-cycle -- from the first persistent level down through the
  single-workgroup tail and back -- against the host restatement built from
  independently assembled level matrices (gpu_helpers.host_cycle): 128x64x64
  (persistent levels 2 and 3, 32 bricks each), 64x32x32 (level 2) and the
  256x128x128 benchmark grid (levels 3 and 4), FP64 and FP32.
* Solves with it and with a launch per step (tk_options.persistent_coarse = -1)
  agree: the same iteration counts within one and the same displacements to
  the solve's tolerance, FP64 and mixed, several solves in a row (captured
  chunks replayed), warm-started as the optimizer does.
"""

import numpy as np
import pytest
import torch

from gpu_helpers import dev, host_cycle, level_stack, make_plan, note, to_dev
from pytopo3d_amd import edge_load_force_vector

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nelx,nely,nelz,first,count", [(128, 64, 64, 2, 2), (64, 32, 32, 2, 1),
                                                        (256, 128, 128, 3, 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_cycle_matches_host_cycle(nelx, nely, nelz, first, count, precision):
    note("persistent cycle", nelx, nely, nelz, precision)
    P, KE = make_plan(nelx, nely, nelz, 1.5)
    levels = [tuple(d) for d in P.levels()]
    note("  levels", levels, "tail", P.tail_level(), "persistent", P.persistent_levels())
    assert P.persistent_levels() == (first, count)
    assert P.tail_level() == first + count == len(levels) - 2
    rho = np.random.default_rng(31).uniform(0.01, 1.0, nelx * nely * nelz)
    P.prepare(to_dev(rho), 3.0, precision=precision)
    lv = level_stack(nelx, nely, nelz, rho, KE, first, len(levels))
    assert [d["dims"] for d in lv] == levels[first:]
    n = lv[0]["A"].shape[0]
    b = np.random.default_rng(6).standard_normal(n)
    b[lv[0]["fixed"]] = 0.0
    want = host_cycle(lv, 0, b, P.galerkin_info()["bound"], P.options)
    dt = torch.float64 if precision == "fp64" else torch.float32
    x = torch.full((n,), np.nan, dtype=dt, device=dev())
    P.coarse_cycle_apply(to_dev(b, dt), x, precision=precision)
    got = x.double().cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    note(f"  {n} DOFs from level {first}: rel err {err:.2e}")
    assert err <= (1e-12 if precision == "fp64" else 1e-4)
    assert np.all(got[lv[0]["fixed"]] == 0.0)
    # the same twice in a row (barrier words and buffers left as they should be)
    x2 = torch.full((n,), np.nan, dtype=dt, device=dev())
    P.coarse_cycle_apply(to_dev(b, dt), x2, precision=precision)
    assert torch.equal(x, x2)
    P.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_cycle_solves_agree_with_launch_per_step(precision):
    nelx, nely, nelz = 128, 64, 64
    note("persistent vs per-step solves", nelx, nely, nelz, precision)
    Fh = edge_load_force_vector(nelx, nely, nelz)
    F = to_dev(Fh)
    rng = np.random.default_rng(8)
    designs = [rng.uniform(0.05, 1.0, nelx * nely * nelz) for _ in range(3)]
    rtol = 1e-10 if precision == "fp64" else 1e-8
    out = {}
    for pc in (1, -1):
        P, _ = make_plan(nelx, nely, nelz, 1.5, options={"persistent_coarse": pc})
        assert (P.persistent_levels()[1] > 0) == (pc == 1)
        U = torch.zeros(P.ndof, dtype=torch.float64, device=dev())
        res = []
        for k, rho in enumerate(designs):
            it, rr = P.solve(to_dev(rho), 3.0, F, U, rtol=rtol, precision=precision,
                             warm_start=k > 0)
            res.append((it, U.cpu().numpy().copy()))
            note(f"  persistent {pc} solve {k}: {it} iterations, relres {rr:.2e}")
        out[pc] = res
        P.close()
    for (it1, u1), (it0, u0) in zip(out[1], out[-1]):
        d = float(np.abs(u1 - u0).max() / np.abs(u0).max())
        dc = abs(float(Fh @ u1) / float(Fh @ u0) - 1.0)
        note(f"  iterations {it1} / {it0}, displacement difference {d:.2e}, compliance {dc:.2e}")
        assert abs(it1 - it0) <= 1
        assert d <= (1e-8 if precision == "fp64" else 1e-6)
        assert dc <= (1e-10 if precision == "fp64" else 1e-8)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_expired_barrier_wait_falls_back_to_a_launch_per_step(precision):
    """A barrier wait limit of zero (tk_plan_set_barrier_timeout) makes the
    persistent cycle's own bounded waits expire at once, so every workgroup
    still exits: the solve fails with the timeout message, the plan drops its
    persistent levels, and the next solve runs them a launch per step, as a
    plan made that way does."""
    nelx, nely, nelz = 128, 64, 64
    note("expired barrier wait", nelx, nely, nelz, precision)
    F = to_dev(edge_load_force_vector(nelx, nely, nelz))
    rho = to_dev(np.random.default_rng(9).uniform(0.05, 1.0, nelx * nely * nelz))
    rtol = 1e-10 if precision == "fp64" else 1e-8
    P, _ = make_plan(nelx, nely, nelz, 1.5)
    assert P.persistent_levels()[1] > 0
    P.set_barrier_timeout(0.0)
    U = torch.zeros(P.ndof, dtype=torch.float64, device=dev())
    with pytest.raises(RuntimeError, match="timed out"):
        P.solve(rho, 3.0, F, U, rtol=rtol, precision=precision)
    assert P.persistent_levels() == (-1, 0)
    U.zero_()
    it, rr = P.solve(rho, 3.0, F, U, rtol=rtol, precision=precision)
    note(f"  after the fallback: {it} iterations, relres {rr:.2e}")
    Q, _ = make_plan(nelx, nely, nelz, 1.5, options={"persistent_coarse": -1})
    V = torch.zeros(Q.ndof, dtype=torch.float64, device=dev())
    it0, _ = Q.solve(rho, 3.0, F, V, rtol=rtol, precision=precision)
    assert it == it0
    d = float((U - V).abs().max() / V.abs().max())
    note(f"  against a launch-per-step plan: {it0} iterations, displacement difference {d:.2e}")
    assert d <= 1e-14
    with pytest.raises(ValueError):
        P.set_barrier_timeout(-1.0)
    P.close()
    Q.close()
