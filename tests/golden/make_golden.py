"""Write the golden trajectory fixtures for ``top3d`` from the reference itself.

Run in the development container only (the reference tree is not present on
GPU machines):

    MPLBACKEND=Agg python tests/golden/make_golden.py

Only ``pytopo3d.core.optimizer.top3d`` is run: the full design trajectory
(per-iteration compliance, the physical density snapshots, stage summaries and
final diagnostics) is the one part of the path that no closed form, invariant or
brute-force check fixes.  Everything the fixtures hold is data: the inputs of
each case and the reference's outputs.  Nothing from the reference's source is
stored.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

REFERENCE = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    # 1) reference default BC (z line load at x = nelx, z = 0; x = 0 face fixed)
    yield "default_bc_6x4x3", dict(
        nelx=6, nely=4, nelz=3, volfrac=0.5, penal=3.0, rmin=1.5, disp_thres=0.5,
        elem_size=1.0, tolx=1e-12, maxloop=3, beta_schedule=(1.0, 2.0),
        save_history=True, history_frequency=1)
    # 2) element force field + support mask + obstacle, default elem_size 0.01
    nelx, nely, nelz = 5, 4, 3
    shape = (nely, nelx, nelz)
    ff = np.zeros(shape + (3,))
    ff[0, nelx - 1, :, 1] = -1.0 / nelz
    sm = np.zeros(shape, bool)
    sm[:, 0, :] = True
    ob = np.zeros(shape, bool)
    ob[2, 2, 1] = True
    yield "force_field_5x4x3", dict(
        nelx=nelx, nely=nely, nelz=nelz, volfrac=0.45, penal=3.0, rmin=1.6, disp_thres=0.5,
        force_field=ff, support_mask=sm, obstacle_mask=ob, tolx=1e-12, maxloop=2,
        beta_schedule=(1.0, 4.0), move=0.15, save_history=True, history_frequency=1)
    # 3) orthotropic material, protected zone mask, projection eta 0.4
    nelx, nely, nelz = 6, 3, 2
    pz = np.zeros((nely, nelx, nelz), bool)
    pz[1, 3, :] = True
    yield "orthotropic_6x3x2", dict(
        nelx=nelx, nely=nely, nelz=nelz, volfrac=0.55, penal=3.0, rmin=2.0, disp_thres=0.5,
        material_params=(10.0, 20.0, 30.0, 4.0, 5.0, 6.0, 0.10, 0.12, 0.08), elem_size=0.5,
        protected_zone_mask=pz, tolx=1e-12, maxloop=3, beta_schedule=(2.0,),
        projection_eta=0.4, save_history=True, history_frequency=1)


def main():
    sys.path.insert(0, REFERENCE)
    from pytopo3d.core.optimizer import top3d  # noqa: E402  (reference, this container only)

    index = {}
    for name, kw in _cases():
        diag = {}
        rho, hist, c, _ = top3d(diagnostics_out=diag, **kw)
        arrays = {
            "rho": np.asarray(rho),
            "density_history": np.stack(hist["density_history"]),
            "iteration_history": np.asarray(hist["iteration_history"]),
            "compliance_history": np.asarray(hist["compliance_history"]),
            "beta_history": np.asarray(hist["beta_history"]),
        }
        inputs = {}
        for key, val in kw.items():
            if isinstance(val, np.ndarray):
                arrays["in_" + key] = val
            else:
                inputs[key] = list(val) if isinstance(val, tuple) else val
        np.savez_compressed(os.path.join(HERE, f"top3d_{name}.npz"), **arrays)
        index[name] = {"inputs": inputs, "final_compliance": float(c), "diagnostics": diag}
        print(name, "c =", c, "iters =", diag["projection_total_iterations"])
    with open(os.path.join(HERE, "top3d_golden.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
