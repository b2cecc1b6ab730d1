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


_C_CALLER = r"""
#include "pytopo3d_hip.h"
#include <stdio.h>
int main(void) {
  tk_options o;
  tk_options_default(&o);
  /* calls that fail in argument checking, before any device work */
  int rc = tk_plan_create(NULL, 4, 4, 4, NULL, 1.5, &o);
  printf("%d %d %d %s\n", o.smooth_degree, o.coarse_degree, rc, tk_last_error());
  static double KE[576];
  for (int i = 0; i < 24; ++i) KE[i * 25] = 1.0;
  tk_plan* p = NULL;
  o.mid_cheb_lo = 2.0;   /* above cheb_hi */
  rc = tk_plan_create(&p, 4, 4, 4, KE, 1.5, &o);
  printf("| %d %d %s\n", rc, p == NULL, tk_last_error());
  tk_options_default(&o);
  o.direction_slots = 1;
  rc = tk_plan_create(&p, 4, 4, 4, KE, 1.5, &o);
  printf("| %d %d %s\n", rc, p == NULL, tk_last_error());
  return 0;
}
"""


def test_header_is_plain_c_and_links_from_c(tmp_path):
    """A C99 caller includes the header and links the library (no device call)."""
    lib = os.path.join(ROOT, "pytopo3d_amd", "libpytopo3d_hip.so")
    if not os.path.exists(lib):
        from pytopo3d_amd import build
        build.build(verbose=False)
    src = tmp_path / "caller.c"
    src.write_text(_C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(lib)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                    "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lpytopo3d_hip", "-Wl,-rpath," + libdir],
                   check=True, capture_output=True, text=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True,
                           text=True).stdout.split("| ")
    out = lines[0].split()
    assert out[:3] == ["2", "12", "1"]
    assert "out must not be NULL" in " ".join(out[3:])
    # option checks (smoothing intervals, direction slots) refuse before any allocation
    assert lines[1].split()[:2] == ["1", "1"]
    assert "mid_cheb_lo must be below cheb_hi" in lines[1]
    assert lines[2].split()[:2] == ["1", "1"]
    assert "direction_slots must be 0 or in [2, " in lines[2]


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


def test_runner_signature_matches_the_reference_runner():
    """execute_optimization takes the reference runner's parameters in its
    order with its defaults (pytopo3d/runners/experiment.py:95-119)."""
    import inspect

    from pytopo3d_amd.runner import execute_optimization
    want = [("nelx", None), ("nely", None), ("nelz", None), ("volfrac", None), ("penal", None),
            ("rmin", None), ("disp_thres", None), ("material_params", None),
            ("elem_size", 0.01), ("force_field", None), ("support_mask", None), ("tolx", 0.01),
            ("maxloop", 2000), ("create_animation", False), ("animation_frequency", 10),
            ("logger", None), ("combined_obstacle_mask", None), ("use_gpu", False),
            ("protected_zone_mask", None), ("beta_schedule", (1.0, 2.0, 4.0, 8.0)),
            ("projection_eta", 0.5), ("move", 0.2), ("diagnostics_out", None)]
    params = [p for p in inspect.signature(execute_optimization).parameters.values()
              if p.kind is not inspect.Parameter.VAR_KEYWORD]
    assert [p.name for p in params] == [n for n, _ in want]
    for p, (_, d) in zip(params, want):
        if p.default is not inspect.Parameter.empty:
            assert p.default == d, p.name


def test_preset_helpers_match_reference_semantics():
    """Preset boxes of pytopo3d/utils/assembly.py:28-74, :186-222 and the
    protected-zone ranges of main.py:158-168."""
    from pytopo3d_amd.problem import (build_force_field, build_support_mask,
                                      protected_zone_mask_from_ranges)
    n = (12, 6, 4)
    ff = build_force_field(*n, 10, 12, 0, 6, 0, 4, 1.0, -2.0, 0.0)
    assert ff.shape == (6, 12, 4, 3)
    np.testing.assert_allclose(ff.reshape(-1, 3).sum(axis=0), [1.0, -2.0, 0.0])
    assert np.count_nonzero(ff[..., 0]) == 2 * 6 * 4
    with pytest.raises(ValueError, match="out of bounds"):
        build_force_field(*n, 10, 13, 0, 6, 0, 4, 1.0, 0.0, 0.0)
    with pytest.raises(ValueError, match="out of bounds"):
        build_support_mask(*n, 3, 3, 0, 6, 0, 4)
    m = protected_zone_mask_from_ranges(*n, [(0, 2, 0, 6, 0, 4), (1, 3, 2, 3, 1, 2)])
    assert m.sum() == 2 * 6 * 4 + 1


def test_galerkin_pattern_header_matches_its_generator():
    """tk_galerkin_pattern.hpp is what gen_galerkin.py derives from lk_H8 (the
    Galerkin coupling pattern, checked there on three orthotropic KEs)."""
    import importlib.util
    path = os.path.join(ROOT, "pytopo3d_amd", "csrc", "gen_galerkin.py")
    spec = importlib.util.spec_from_file_location("gen_galerkin", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    pat = gen.pattern()
    with open(os.path.join(ROOT, "pytopo3d_amd", "csrc", "tk_galerkin_pattern.hpp")) as fh:
        text = fh.read()

    def arr(name):
        m = re.search(name + r"\[[^\]]*\] = \{([^}]*)\}", text)
        return [int(v) for v in m.group(1).split(",")]

    ws, rows, start, cols = arr("kGalW"), arr("kGalRow"), arr("kGalStart"), arr("kGalCol")
    got = sorted((ws[g], rows[g], cols[k]) for g in range(len(ws))
                 for k in range(start[g], start[g + 1]))
    assert got == pat and len(pat) == 279
    # the straight-line code applies exactly these entries, in table order
    terms = re.findall(r"gk\[(\d+)\] \* a\[(\d)\]\[(\d)\]", text)
    assert [int(k) for k, _, _ in terms] == list(range(279))
    assert [3 * int(m) + int(c) for _, c, m in terms] == cols
