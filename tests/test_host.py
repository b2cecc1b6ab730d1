"""Host-side logic of the product (CPU): problem set-up against the oracle,
argument validation, and the C ABI library's exported symbols."""

import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from pytopo3d_amd import _native
from pytopo3d_amd import problem as pb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORTHO = (10.0, 20.0, 30.0, 4.0, 5.0, 6.0, 0.10, 0.12, 0.08)


@pytest.mark.parametrize("params,h", [(None, 1.0), (None, 0.01), (ORTHO, 0.7)])
def test_element_stiffness_bitwise_equal_to_oracle(params, h):
    assert np.array_equal(pb.element_stiffness(params, h), oracle.element_stiffness(params, h))


def test_force_vectors_bitwise_equal_to_oracle():
    rng = np.random.default_rng(0)
    ff = rng.standard_normal((4, 5, 3, 3)) / 3.0
    ff[rng.random((4, 5, 3)) < 0.4] = 0.0
    assert np.array_equal(pb.build_force_vector(5, 4, 3, ff), oracle.force_vector(5, 4, 3, ff))
    assert np.array_equal(pb.build_force_vector(5, 4, 3), oracle.force_vector(5, 4, 3))
    assert np.array_equal(pb.edge_load_force_vector(6, 3, 4),
                          oracle.edge_load_force_vector(6, 3, 4))
    with pytest.raises(ValueError):
        pb.build_force_vector(5, 4, 3, np.zeros((4, 5, 2, 3)))


@pytest.mark.parametrize("density", [0.0, 0.1, 0.5])
def test_support_flags_match_oracle_fixed_dofs(density):
    rng = np.random.default_rng(1)
    sm = rng.random((4, 5, 3)) < density
    flags = pb.support_node_flags(5, 4, 3, sm)
    _, fixed = oracle.support_dofs(5, 4, 3, sm)
    np.testing.assert_array_equal(np.flatnonzero(np.repeat(flags > 0, 3)), fixed)


def test_protected_masks_match_oracle():
    rng = np.random.default_rng(2)
    shape = (4, 5, 3)
    ob = rng.random(shape) < 0.2
    pz = rng.random(shape) < 0.2
    ff = np.zeros(shape + (3,)); ff[0, 4, :, 1] = -1.0
    sm = np.zeros(shape, bool); sm[:, 0, :] = True
    for args in [(ob, pz, None, None, None), (ob, pz, ff, sm, None), (None, None, None, None, None),
                 (ob, None, None, None, pb.edge_load_force_vector(5, 4, 3))]:
        a = pb.protected_masks(5, 4, 3, *args)
        b = oracle.boundary_protection(5, 4, 3, *args)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)


def test_edge_load_protects_the_loaded_edge_strip():
    solid, void, free = pb.protected_masks(6, 3, 4, force_vector=pb.edge_load_force_vector(6, 3, 4))
    want = np.zeros((3, 6, 4), bool)
    want[:, 0, :] = True
    want[0, 5, :] = True
    np.testing.assert_array_equal(solid, want)


def test_element_flags_fortran_order():
    solid = np.zeros((2, 3, 2), bool); solid[1, 2, 0] = True
    void = np.zeros((2, 3, 2), bool); void[0, 0, 1] = True
    f = pb.element_flags(solid, void)
    assert f[1 + 2 * 2 + 0] == pb.SOLID and f[0 + 0 + 1 * 6] == pb.VOID and f.sum() == 3


def _header_symbols():
    with open(os.path.join(ROOT, "include", "pytopo3d_hip.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(tk_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "pytopo3d_amd", "libpytopo3d_hip.so")
    if not os.path.exists(lib):
        from pytopo3d_amd import build
        build.build(verbose=False)
    syms = _header_symbols()
    assert len(syms) >= 19
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (tk_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert set(_native.EXPORTED) == set(syms)
    cdll = _native.load_library(lib)         # loads without a GPU
    for s in syms:
        assert hasattr(cdll, s)


def test_product_refuses_to_run_without_a_gpu(monkeypatch):
    import torch
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        _native.lib()


def test_top3d_validates_like_the_reference():
    from pytopo3d_amd import top3d
    with pytest.raises(ValueError, match="volfrac"):
        top3d(2, 2, 2, 0.0, 3.0, 1.5, 0.5)
    with pytest.raises(ValueError, match="penal"):
        top3d(2, 2, 2, 0.5, -1.0, 1.5, 0.5)
    with pytest.raises(ValueError, match="beta_schedule"):
        top3d(2, 2, 2, 0.5, 3.0, 1.5, 0.5, beta_schedule=())
    with pytest.raises(ValueError, match="projection_eta"):
        top3d(2, 2, 2, 0.5, 3.0, 1.5, 0.5, projection_eta=1.0)
    with pytest.raises(ValueError, match="precision"):
        top3d(2, 2, 2, 0.5, 3.0, 1.5, 0.5, precision="fp16")


def test_oc_bisection_with_probes_equals_plain_bisection(tmp_path):
    """The device OC bisection logic (pytopo3d_amd/csrc/tk_bisect.hpp, compiled
    here for the host) against the plain bisection of
    pytopo3d/utils/oc_update.py:76-124 on 1300 seeded monotone volume functions:
    the kept multiplier, its error and the step count must be identical, and
    smooth cases must need far fewer volume evaluations."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc is needed to build the host check")
    exe = str(tmp_path / "oc_bisect_check")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(ROOT, "pytopo3d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "oc_bisect_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    smooth = [ln for ln in res.stdout.splitlines() if ln.strip().startswith("smooth")]
    assert smooth
    for ln in smooth:
        plain, probes = map(float, re.findall(r"plain ([\d.]+) probes ([\d.]+)", ln)[0])
        assert probes < 0.3 * plain, ln
