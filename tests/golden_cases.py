"""Loaders of the reference-written fixtures under tests/golden/ shared by the
oracle (CPU) and HIP (GPU) tests."""

import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _metrics_cases():
    with open(os.path.join(GOLDEN, "fixed_geometry_metrics_golden.json")) as fh:
        index = json.load(fh)
    for name, case in sorted(index.items()):
        x = np.asarray(case["design"]).reshape(case["shape"], order="F")
        kw = dict(case["inputs"])
        if "material_params" in kw:
            kw["material_params"] = tuple(kw["material_params"])
        if case["arrays_from"]:
            z = np.load(os.path.join(GOLDEN, case["arrays_from"]))
            for k in ("force_field", "support_mask", "obstacle_mask", "protected_zone_mask"):
                if "in_" + k in z.files:
                    kw[k] = z["in_" + k]
        else:   # the two-component field case of tests/golden/make_metrics_golden.py
            ff = np.zeros(tuple(case["shape"]) + (3,))
            ff[0, 5, :, 1] = -0.5
            ff[1, 5, :, 2] = 0.25
            kw["force_field"] = ff
        yield name, x, kw, case["metrics"]


METRIC_CASES = list(_metrics_cases())


def check_metrics(got, want):
    """Every key of the reference's metrics dict: None where it is None, the
    counts exactly, floats to 1e-9 relative (absolute 1e-9 of the largest
    displacement average for the averages that vanish by symmetry)."""
    assert set(got) == set(want)
    uscale = max(abs(want[k]) for k in ("ux_avg_load_patch", "uy_avg_load_patch",
                                        "uz_avg_load_patch") if want[k] is not None)
    for k, v in want.items():
        if v is None:
            assert got[k] is None, k
        elif isinstance(v, int) and not isinstance(v, bool):
            assert got[k] == v, k
        elif k.endswith("_avg_load_patch"):
            assert got[k] == pytest.approx(v, rel=1e-9, abs=1e-9 * uscale), k
        else:
            assert got[k] == pytest.approx(v, rel=1e-9), k


