"""ctypes binding of libpytopo3d_hip.so (the C ABI in include/pytopo3d_hip.h).

Device buffers are torch tensors on the current HIP device; their data
pointers and torch's current stream are handed to the library.  There is no
fallback: if the library or a GPU is missing, :func:`lib` raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpytopo3d_hip.so")

TK_FP64 = 0
TK_MIXED_FP32 = 1
TK_SOLVE_X0 = 1
TK_SOLVE_JACOBI = 2
TK_SOLVE_REUSE = 4

# "mixed": FP64 flexible CG around an FP32 multigrid correction.  There is no
# all-FP32 solve: the FP64 outer iteration is what reaches FP64 accuracy at the
# 1e-9 SIMP contrast.
PRECISIONS = {"fp64": TK_FP64, "mixed": TK_MIXED_FP32, "mixed_fp32": TK_MIXED_FP32}


class SolveNotConverged(RuntimeError):
    """The matrix-free Krylov solve missed its tolerance (the reference raises
    RuntimeError("CG failed to converge ..."), pytopo3d/utils/solver.py:57-58)."""


class TkOptions(ctypes.Structure):
    _fields_ = [
        ("smooth_degree", ctypes.c_int),
        ("cheb_lo", ctypes.c_double),
        ("cheb_hi", ctypes.c_double),
        ("max_levels", ctypes.c_int),
        ("coarse_max_dof", ctypes.c_int),
        ("check_every", ctypes.c_int),
        ("operator_kind", ctypes.c_int),
        ("no_graphs", ctypes.c_int),
        ("filter_kind", ctypes.c_int),
        ("coarse_degree", ctypes.c_int),
        ("coarse_cheb_lo", ctypes.c_double),
        ("direction_slots", ctypes.c_int),
        ("galerkin", ctypes.c_int),
        ("persistent_coarse", ctypes.c_int),
        ("mid_degree", ctypes.c_int),
        ("mid_cheb_lo", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SIGS = {
    "tk_options_default": (None, [ctypes.POINTER(TkOptions)]),
    "tk_plan_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, ctypes.POINTER(_D), _D,
                            ctypes.POINTER(TkOptions)]),
    "tk_plan_destroy": (_I, [_P]),
    "tk_plan_set_masks": (_I, [_P, _P, _P, ctypes.POINTER(ctypes.c_longlong), _P]),
    "tk_build_boundary_conditions": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tk_plan_levels": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), _I]),
    "tk_plan_stencil": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_D), _I, ctypes.POINTER(_I)]),
    "tk_plan_filter_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "tk_plan_galerkin_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "tk_level_apply": (_I, [_P, _I, _I, _P, _P, _P]),
    "tk_tail_apply": (_I, [_P, _I, _P, _P, _P]),
    "tk_plan_tail_level": (_I, [_P, ctypes.POINTER(_I)]),
    "tk_plan_persistent_levels": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "tk_coarse_cycle_apply": (_I, [_P, _I, _P, _P, _P]),
    "tk_transfer_apply": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "tk_plan_set_barrier_timeout": (_I, [_P, _D]),
    "tk_filter_row_sums": (_I, [_P, _P, _P]),
    "tk_filter_project": (_I, [_P, _P, _D, _D, _P, _P, _P, _P]),
    "tk_filter_transpose2": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "tk_apply_K": (_I, [_P, _I, _P, _D, _P, _P, _P]),
    "tk_diag_K": (_I, [_P, _P, _D, _P, _P]),
    "tk_smoother_step": (_I, [_P, _I, _P, _P, _P, _P, _D, _D, _P]),
    "tk_prepare": (_I, [_P, _I, _P, _D, _P]),
    "tk_solve": (_I, [_P, _I, _P, _D, _P, _P, _D, _I, _I, ctypes.POINTER(_I),
                      ctypes.POINTER(_D), _P]),
    "tk_energy": (_I, [_P, _P, _P, _P, _D, _P, _P, _P, _P, ctypes.POINTER(_D), _P]),
    "tk_oc_update": (_I, [_P, _P, _P, _P, _D, _D, _D, _D, _I, _D, _P, ctypes.POINTER(_I), _P]),
    "tk_design_stats": (_I, [_P, _P, _P, _P, _P]),
    "tk_volume_bounds": (_I, [_P, _D, _D, ctypes.POINTER(_D), ctypes.POINTER(_D), _P]),
    "tk_read_results": (_I, [_P, ctypes.POINTER(_D), _I, _P]),
    "tk_plan_counters": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _I]),
    "tk_energy_noalloc": (_I, [_I, _I, _I, ctypes.POINTER(_D), _P, _P, _D, _P, _P, _P,
                               ctypes.POINTER(_D), _P]),
    "tk_last_error": (ctypes.c_char_p, []),
}
EXPORTED = tuple(_SIGS)

_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the shared library and declare every entry point (no GPU needed)."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m pytopo3d_amd.build` "
            "(there is no CPU fallback)")
    cdll = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(cdll, name)
        fn.restype = res
        fn.argtypes = args
    return cdll


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not torch.cuda.is_available():
            raise RuntimeError("pytopo3d_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _lib = load_library()
    return _lib


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = (lib().tk_last_error() or b"").decode(errors="replace")
    if rc == 1:
        raise ValueError(msg)
    if rc == 3:
        raise SolveNotConverged(msg)
    raise RuntimeError(msg)


def _ptr(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("device tensors must be contiguous CUDA/HIP tensors")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Plan:
    """One problem size: stencil filter, masks, multigrid hierarchy, workspaces."""

    def __init__(self, nelx: int, nely: int, nelz: int, KE: np.ndarray, rmin: float,
                 options: Optional[dict] = None):
        L = lib()
        ke = np.ascontiguousarray(np.asarray(KE, dtype=np.float64).reshape(24, 24))
        opt = TkOptions()
        L.tk_options_default(ctypes.byref(opt))
        for k, v in (options or {}).items():
            if not hasattr(opt, k):
                raise ValueError(f"unknown solver option {k!r}")
            setattr(opt, k, v)
        h = _P()
        _check(L.tk_plan_create(ctypes.byref(h), int(nelx), int(nely), int(nelz),
                                ke.ctypes.data_as(ctypes.POINTER(_D)), float(rmin),
                                ctypes.byref(opt)))
        self._h = h
        self.nelx, self.nely, self.nelz = int(nelx), int(nely), int(nelz)
        self.nele = self.nelx * self.nely * self.nelz
        self.nnode = (self.nelx + 1) * (self.nely + 1) * (self.nelz + 1)
        self.ndof = 3 * self.nnode
        self.nfree = self.nele
        self.options = {k: getattr(opt, k) for k, _ in TkOptions._fields_}

    # -- lifecycle -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tk_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ----------------------------------------------------------
    def set_masks(self, elem_flags: torch.Tensor, node_fixed: torch.Tensor) -> int:
        assert elem_flags.dtype == torch.uint8 and elem_flags.numel() == self.nele
        assert node_fixed.dtype == torch.uint8 and node_fixed.numel() == self.nnode
        n = ctypes.c_longlong()
        _check(lib().tk_plan_set_masks(self._h, _ptr(elem_flags), _ptr(node_fixed),
                                       ctypes.byref(n), _stream()))
        self.nfree = int(n.value)
        return self.nfree

    def boundary_conditions(self, force_field=None, support_mask=None, obstacle_mask=None,
                            protected_zone_mask=None, F=None, node_fixed=None, elem_flags=None):
        """tk_build_boundary_conditions: inputs are device tensors in the
        reference's layouts (C-order (nely, nelx, nelz[, 3])), any may be None."""
        _check(lib().tk_build_boundary_conditions(
            self._h, _ptr(force_field), _ptr(support_mask), _ptr(obstacle_mask),
            _ptr(protected_zone_mask), _ptr(F), _ptr(node_fixed), _ptr(elem_flags), _stream()))

    def levels(self):
        n = _I()
        dims = (_I * 36)()
        _check(lib().tk_plan_levels(self._h, ctypes.byref(n), dims, 12))
        return [tuple(dims[3 * i:3 * i + 3]) for i in range(n.value)]

    def stencil(self):
        n = _I()
        _check(lib().tk_plan_stencil(self._h, None, None, 0, ctypes.byref(n)))
        off = np.zeros((n.value, 3), dtype=np.int32)
        w = np.zeros(n.value)
        _check(lib().tk_plan_stencil(self._h, off.ctypes.data_as(ctypes.POINTER(_I)),
                                     w.ctypes.data_as(ctypes.POINTER(_D)), n.value,
                                     ctypes.byref(n)))
        return off, w   # columns (dy, dx, dz)

    def filter_info(self) -> dict:
        """Filter kernel of the plan: {"tiled": bool, "halo": int, "nsten": int}."""
        t, h, n = _I(), _I(), _I()
        _check(lib().tk_plan_filter_info(self._h, ctypes.byref(t), ctypes.byref(h),
                                         ctypes.byref(n)))
        return {"tiled": bool(t.value), "halo": int(h.value), "nsten": int(n.value)}

    def galerkin_info(self) -> dict:
        """{"level": the Galerkin level (-1: none), "bound": its lambda_max(D^-1 A) bound}."""
        lv, b = _I(), _D()
        _check(lib().tk_plan_galerkin_info(self._h, ctypes.byref(lv), ctypes.byref(b)))
        return {"level": int(lv.value), "bound": float(b.value)}

    def level_apply(self, level: int, u, y, precision="fp64"):
        """y = A_level u with the hierarchy of the last solve/prepare in this precision."""
        _check(lib().tk_level_apply(self._h, PRECISIONS[precision], int(level), _ptr(u), _ptr(y),
                                    _stream()))

    def tail_level(self) -> int:
        """The level whose V-cycle part runs in one workgroup (-1: none)."""
        lv = _I()
        _check(lib().tk_plan_tail_level(self._h, ctypes.byref(lv)))
        return int(lv.value)

    def persistent_levels(self):
        """(first, count) of the persistent coarse cycle's levels (first -1: none)."""
        f, n = _I(), _I()
        _check(lib().tk_plan_persistent_levels(self._h, ctypes.byref(f), ctypes.byref(n)))
        return int(f.value), int(n.value)

    def transfer_apply(self, level, direction, src, dst, precision="fp64"):
        """dst = R src (direction 0, level -> level + 1) or dst += P src
        (direction 1, level + 1 -> level)."""
        _check(lib().tk_transfer_apply(self._h, PRECISIONS[precision], int(level), int(direction),
                                       _ptr(src), _ptr(dst), _stream()))

    def set_barrier_timeout(self, seconds):
        """Wait limit of the persistent coarse cycle's barriers (default 0.05 s)."""
        _check(lib().tk_plan_set_barrier_timeout(self._h, float(seconds)))

    def coarse_cycle_apply(self, b, x, precision="fp64"):
        """x = the V-cycle part from the first persistent level down and back, for b there."""
        _check(lib().tk_coarse_cycle_apply(self._h, PRECISIONS[precision], _ptr(b), _ptr(x),
                                           _stream()))

    def tail_apply(self, b, x, precision="fp64"):
        """x = the single-workgroup V-cycle tail for right-hand side b on the tail level."""
        _check(lib().tk_tail_apply(self._h, PRECISIONS[precision], _ptr(b), _ptr(x), _stream()))

    def row_sums(self, out: torch.Tensor) -> torch.Tensor:
        _check(lib().tk_filter_row_sums(self._h, _ptr(out), _stream()))
        return out

    # -- hot path --------------------------------------------------------
    def filter_project(self, x, beta, eta, rho_f=None, rho=None, drho=None):
        _check(lib().tk_filter_project(self._h, _ptr(x), float(beta), float(eta), _ptr(rho_f),
                                       _ptr(rho), _ptr(drho), _stream()))

    def filter_transpose2(self, g_c, g_v, drho, dc_dx, dv_dx=None):
        _check(lib().tk_filter_transpose2(self._h, _ptr(g_c), _ptr(g_v), _ptr(drho),
                                          _ptr(dc_dx), _ptr(dv_dx), _stream()))

    def apply_K(self, rho, penal, u, y, precision: str = "fp64"):
        _check(lib().tk_apply_K(self._h, PRECISIONS[precision], _ptr(rho), float(penal),
                                _ptr(u), _ptr(y), _stream()))

    def smoother_step(self, b, x, xo, d, c1, c2, precision="fp64"):
        _check(lib().tk_smoother_step(self._h, PRECISIONS[precision], _ptr(b), _ptr(x), _ptr(xo),
                                      _ptr(d), float(c1), float(c2), _stream()))

    def diag_K(self, rho, penal, diag):
        _check(lib().tk_diag_K(self._h, _ptr(rho), float(penal), _ptr(diag), _stream()))

    def prepare(self, rho, penal, precision="fp64"):
        """Build the multigrid state of a design; solve(..., reuse=True) uses it."""
        _check(lib().tk_prepare(self._h, PRECISIONS[precision], _ptr(rho), float(penal),
                                _stream()))

    def solve(self, rho, penal, F, U, rtol=1e-10, maxit=5000, precision="fp64",
              warm_start=False, jacobi=False, reuse=False):
        flags = (TK_SOLVE_X0 if warm_start else 0) | (TK_SOLVE_JACOBI if jacobi else 0) | (
            TK_SOLVE_REUSE if reuse else 0)
        it = _I()
        rr = _D()
        _check(lib().tk_solve(self._h, PRECISIONS[precision], _ptr(rho), float(penal), _ptr(F),
                              _ptr(U), float(rtol), int(maxit), flags, ctypes.byref(it),
                              ctypes.byref(rr), _stream()))
        return int(it.value), float(rr.value)

    def energy(self, U, rho, penal, drho=None, ce=None, dc_drho=None, dc_dx=None, dv_dx=None,
               sync=True):
        c = _D()
        _check(lib().tk_energy(self._h, _ptr(U), _ptr(rho), _ptr(drho), float(penal), _ptr(ce),
                               _ptr(dc_drho), _ptr(dc_dx), _ptr(dv_dx),
                               ctypes.byref(c) if sync else None, _stream()))
        return float(c.value) if sync else None

    def oc_update(self, x, dc_dx, dv_dx, move, target, beta, eta, x_new, max_steps=80,
                  tol=1e-8, sync=False):
        st = _I()
        _check(lib().tk_oc_update(self._h, _ptr(x), _ptr(dc_dx), _ptr(dv_dx), float(move),
                                  float(target), float(beta), float(eta), int(max_steps),
                                  float(tol), _ptr(x_new), ctypes.byref(st) if sync else None,
                                  _stream()))
        return int(st.value) if sync else None

    def design_stats(self, rho, x_old=None, x_new=None):
        _check(lib().tk_design_stats(self._h, _ptr(x_old), _ptr(x_new), _ptr(rho), _stream()))

    def volume_bounds(self, beta, eta):
        lo, hi = _D(), _D()
        _check(lib().tk_volume_bounds(self._h, float(beta), float(eta), ctypes.byref(lo),
                                      ctypes.byref(hi), _stream()))
        return float(lo.value), float(hi.value)

    def results(self) -> dict:
        out = (_D * 10)()
        _check(lib().tk_read_results(self._h, out, 10, _stream()))
        keys = ("c", "change", "vol_sum", "gray_cnt", "oc_steps", "oc_lambda", "oc_volume",
                "ratio_scale", "nfree", "oc_evals")
        return dict(zip(keys, list(out)))

    def counters(self) -> dict:
        out = (ctypes.c_longlong * 7)()
        _check(lib().tk_plan_counters(self._h, out, 7))
        keys = ("applies64", "applies32", "cg_iters", "ir_steps", "solves", "oc_steps",
                "filter_passes")
        return dict(zip(keys, list(out)))


def energy_noalloc(nelx: int, nely: int, nelz: int, KE: np.ndarray, U: torch.Tensor,
                   rho: torch.Tensor, penal: float, ce: Optional[torch.Tensor] = None,
                   dc_drho: Optional[torch.Tensor] = None, c_dev: Optional[torch.Tensor] = None,
                   sync: bool = True) -> Optional[float]:
    """tk_energy_noalloc: energies and compliance without a plan or any device
    allocation (the library's static scratch holds the reduction).  Returns c
    when ``sync`` (the stream is synchronised), else None (c goes to c_dev)."""
    ke = np.ascontiguousarray(np.asarray(KE, dtype=np.float64).reshape(24, 24))
    c = _D()
    _check(lib().tk_energy_noalloc(int(nelx), int(nely), int(nelz),
                                   ke.ctypes.data_as(ctypes.POINTER(_D)), _ptr(U), _ptr(rho),
                                   float(penal), _ptr(ce), _ptr(dc_drho), _ptr(c_dev),
                                   ctypes.byref(c) if sync else None, _stream()))
    return float(c.value) if sync else None
