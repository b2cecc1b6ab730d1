"""Write fixed-geometry metric fixtures from the reference itself.

Run in the development container only (the reference tree is not present on
GPU machines):

    MPLBACKEND=Agg python tests/golden/make_metrics_golden.py

Calls the reference's ``evaluate_fixed_geometry_metrics``
(pytopo3d/core/optimizer.py:54-252) on the final designs of the three golden
``top3d`` cases (tests/golden/top3d_*.npz) with those cases' material, loads,
supports and masks, plus one design with a two-component element force field,
and stores the returned dicts (data only) in fixed_geometry_metrics_golden.json.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

REFERENCE = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    with open(os.path.join(HERE, "top3d_golden.json")) as fh:
        index = json.load(fh)
    for name, meta in sorted(index.items()):
        z = np.load(os.path.join(HERE, f"top3d_{name}.npz"))
        inp = meta["inputs"]
        kw = dict(penal=inp["penal"], elem_size=inp.get("elem_size", 0.01))
        if "material_params" in inp:
            kw["material_params"] = tuple(inp["material_params"])
        for k in ("force_field", "support_mask", "obstacle_mask", "protected_zone_mask"):
            if "in_" + k in z.files:
                kw[k] = z["in_" + k]
        yield name, z["rho"], kw
    rng = np.random.default_rng(12)
    shape = (4, 6, 3)
    ff = np.zeros(shape + (3,))
    ff[0, 5, :, 1] = -0.5
    ff[1, 5, :, 2] = 0.25
    yield "two_component_field_6x4x3", rng.uniform(0.05, 1.0, shape), dict(
        penal=3.0, elem_size=1.0, force_field=ff)


def main():
    sys.path.insert(0, REFERENCE)
    from pytopo3d.core.optimizer import evaluate_fixed_geometry_metrics  # noqa: E402

    out = {}
    for name, x, kw in _cases():
        m = evaluate_fixed_geometry_metrics(xPhys=x, **kw)
        out[name] = {"design": x.ravel(order="F").tolist(), "shape": list(x.shape),
                     "inputs": {k: v for k, v in kw.items() if not isinstance(v, np.ndarray)},
                     "arrays_from": (f"top3d_{name}.npz" if name != "two_component_field_6x4x3"
                                     else None),
                     "metrics": {k: (None if v is None else (float(v) if isinstance(v, float)
                                                             else int(v)))
                                 for k, v in m.items()}}
        print(name, m["compliance"])
    with open(os.path.join(HERE, "fixed_geometry_metrics_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
