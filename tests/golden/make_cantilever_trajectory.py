"""Write the 60x30x20 cantilever trajectory fixture from the oracle.

    python tests/golden/make_cantilever_trajectory.py      (development container, CPU)

The 60x30x20 unit-cube H8 cantilever -- every DOF on x = 0 fixed, a total
unit force in -y spread evenly over the nodes of the edge x = nelx, y = 0 --
with volfrac 0.3, penal 3, rmin 3, beta 1, 2, 4, 8 (25 iterations per stage,
tolx 1e-12), eta 0.5, move 0.2 and x0 = 0.3, run through ``oracle.top3d``.
A SuperLU factorisation of this system takes minutes per solve (3-D fill), so
every solve is the oracle's multigrid-preconditioned CG on the assembled
matrix to a relative residual of 1e-12 (``oracle.MultigridCG``, pinned
against SuperLU on smaller grids in tests/test_oracle_solve_top3d.py).

Stored (data only): per design iteration the compliance, the free volume
fraction of the projected density the iteration solved with, the change and
three sums of the design x after the iteration (plain, of squares, and
weighted by default_rng(60).uniform(-1, 1, nele)); the full x after
iterations 1, 25, 50 and 100; the final projected density and compliance.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

NELX, NELY, NELZ = 60, 30, 20
KEEP = (1, 25, 50, 100)


def main():
    shape = (NELY, NELX, NELZ)
    F = oracle.edge_load_force_vector(NELX, NELY, NELZ)
    free, _ = oracle.support_dofs(NELX, NELY, NELZ)
    solver = oracle.MultigridCG(NELX, NELY, NELZ, free, rtol=1e-12, coarse_dofs=20000)
    nele = NELX * NELY * NELZ
    w = np.random.default_rng(60).uniform(-1.0, 1.0, nele)
    rows = {k: [] for k in ("c", "vol", "change", "beta", "x_sum", "x_sq", "x_w")}
    keep_x = {}
    t0 = time.time()
    _, _, free_m = oracle.boundary_protection(NELX, NELY, NELZ, None, None, None, None, F)

    def record(d):
        x = d["x_new"].ravel(order="F")
        rows["c"].append(d["c"])
        rows["vol"].append(float(np.mean(np.asarray(d["rho"])[free_m])))
        rows["change"].append(d["change"])
        rows["beta"].append(d["beta"])
        rows["x_sum"].append(float(np.sum(x)))
        rows["x_sq"].append(float(np.sum(x * x)))
        rows["x_w"].append(float(np.dot(w, x)))
        if d["total"] in KEEP:
            keep_x[d["total"]] = x.copy()
        print(f"iter {d['total']:3d} beta {d['beta']:g} c {d['c']:.12e} "
              f"cg {solver.iterations[-1]} t {time.time() - t0:.0f}s", flush=True)

    rho, _, c, _ = oracle.top3d(
        NELX, NELY, NELZ, 0.3, 3.0, 3.0, 0.5, elem_size=1.0, tolx=1e-12, maxloop=25,
        beta_schedule=(1.0, 2.0, 4.0, 8.0), projection_eta=0.5, move=0.2,
        force_vector_override=F, solver=solver, record=record)
    out = {k: np.asarray(v) for k, v in rows.items()}
    out["keep_iters"] = np.asarray(KEEP)
    out["keep_x"] = np.stack([keep_x[k] for k in KEEP])
    out["final_rho"] = np.asarray(rho).ravel(order="F")
    out["final_c"] = np.asarray(c)
    out["cg_iterations"] = np.asarray(solver.iterations)
    np.savez_compressed(os.path.join(HERE, "cantilever_60x30x20_trajectory.npz"), **out)
    print("final c", c, "total", time.time() - t0, "s")


if __name__ == "__main__":
    main()
