"""ctypes binding of include/springsteel_hip.h.

This is the Python stand-in for the Julia ``ccall`` shim: the same C entry
points, the same column-major arrays.  The product path has no fallback: when
the HIP library is not built, loading raises ``ExtensionMissingError``.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

import numpy as np
from typing import Optional

from .errors import (ArgumentError, BoundsError, DomainError, ExtensionMissingError, HIPError,
                     NotSupportedError, PosDefException, SpringsteelError)
from .params import ResolvedGrid

LIB_NAME = "libspringsteel_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

SSG_OK = 0
SSG_EINVAL, SSG_ENOTSUP, SSG_EPOSDEF, SSG_EDOMAIN, SSG_EBOUNDS, SSG_ENOMEM, SSG_EHIP = -22, -95, -33, -34, -35, -12, -5
SSG_MEM_HOST, SSG_MEM_DEVICE = 0, 1
SSG_FP64, SSG_FP32_REFINE = 0, 1

# every symbol include/springsteel_hip.h declares
EXPORTED = ("ssg_plan_create", "ssg_plan_destroy", "ssg_plan_get_dims", "ssg_spectral_transform",
            "ssg_grid_transform", "ssg_plan_set_timing", "ssg_plan_stage_times", "ssg_last_error", "ssg_batch_create",
            "ssg_batch_spectral_transform", "ssg_batch_grid_transform", "ssg_batch_tune", "ssg_batch_launches",
            "ssg_batch_destroy", "ssg_plan_get_tuning", "ssg_plan_set_tuning", "ssg_plan_tune",
            "ssg_regular_transform", "ssg_evaluate_unstructured", "ssg_plan_set_ahat", "ssg_plan_set_filter",
            "ssg_apply_filter", "ssg_grid_transform_noalloc", "ssg_spectral_transform_f32",
            "ssg_grid_transform_f32", "ssg_spline_transform", "ssg_tile_transform", "ssg_sum_spectral_tile",
            "ssg_set_spectral_tile", "ssg_sum_shared_spectral")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)


class SsgDesc(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("num_cells", ctypes.c_int), ("mubar", ctypes.c_int),
                ("quadrature", ctypes.c_int), ("iMin", ctypes.c_double), ("iMax", ctypes.c_double),
                ("nvars", ctypes.c_int), ("l_q", _c_double_p), ("bcl_kind", _c_int_p), ("bcl_ab", _c_double_p),
                ("bcr_kind", _c_int_p), ("bcr_ab", _c_double_p), ("kMin", ctypes.c_double),
                ("kMax", ctypes.c_double), ("kDim", ctypes.c_int), ("b_kDim", ctypes.c_int),
                ("bcb_kind", _c_int_p), ("bct_kind", _c_int_p), ("max_wavenumber", _c_int_p),
                ("ring_kmax", ctypes.c_int), ("patchOffsetL", ctypes.c_int), ("precision", ctypes.c_int),
                ("filter_flags", ctypes.c_int), ("wavenumber_ahat", ctypes.c_int)]


class SsgDims(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_long), ("nspec", ctypes.c_long), ("nvars", ctypes.c_int),
                ("nslots", ctypes.c_int), ("iDim", ctypes.c_int), ("b_iDim", ctypes.c_int), ("jDim", ctypes.c_int),
                ("kDim", ctypes.c_int), ("b_kDim", ctypes.c_int), ("kDim_wn", ctypes.c_int)]


_lib: Optional[ctypes.CDLL] = None


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Load the HIP library (raises ExtensionMissingError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ExtensionMissingError(f"{path} is not built; run __graft_entry__.build()")
    # torch ships its own HIP runtime.  Two runtimes share a process only if
    # torch's has enumerated the devices first: started after this library's,
    # it reports no GPU.  Counting devices is enough and creates no context.
    if importlib.util.find_spec("torch") is not None:
        import torch
        torch.cuda.device_count()
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise ExtensionMissingError(f"cannot load {path}: {e}") from e
    lib.ssg_plan_create.argtypes = [ctypes.POINTER(SsgDesc), ctypes.POINTER(ctypes.c_void_p)]
    lib.ssg_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.ssg_plan_get_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(SsgDims)]
    for fn in (lib.ssg_spectral_transform, lib.ssg_grid_transform, lib.ssg_spectral_transform_f32,
               lib.ssg_grid_transform_f32):
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.ssg_grid_transform_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ssg_plan_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ssg_plan_stage_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), _c_double_p,
                                         ctypes.c_int]
    lib.ssg_last_error.restype = ctypes.c_char_p
    lib.ssg_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
    for fn in (lib.ssg_batch_spectral_transform, lib.ssg_batch_grid_transform, lib.ssg_batch_tune):
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.ssg_batch_launches.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.ssg_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.ssg_plan_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.ssg_plan_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.ssg_plan_tune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ssg_regular_transform.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _c_double_p, ctypes.c_int, _c_double_p,
                                          ctypes.c_int, _c_double_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p]
    lib.ssg_evaluate_unstructured.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _c_double_p, ctypes.c_long,
                                              ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_void_p]
    lib.ssg_plan_set_ahat.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.ssg_plan_set_filter.argtypes = [ctypes.c_void_p, _c_double_p, _c_double_p]
    lib.ssg_apply_filter.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    for name in ("ssg_spline_transform", "ssg_tile_transform"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    for name in ("ssg_sum_spectral_tile", "ssg_set_spectral_tile"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.ssg_sum_shared_spectral.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    for name in EXPORTED:
        if name != "ssg_last_error":
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc: int):
    if rc == SSG_OK:
        return
    msg = (load().ssg_last_error() or b"").decode("utf-8", "replace")
    exc = {SSG_EINVAL: ArgumentError, SSG_ENOTSUP: NotSupportedError, SSG_EPOSDEF: PosDefException,
           SSG_EDOMAIN: DomainError, SSG_EBOUNDS: BoundsError, SSG_EHIP: HIPError,
           SSG_ENOMEM: HIPError}.get(rc, SpringsteelError)
    raise exc(f"[{rc}] {msg}")


def make_desc(rg: ResolvedGrid, precision: int = SSG_FP64):
    """Build the ssg_desc and keep the per-variable arrays alive next to it."""
    n = rg.nvars
    keep = dict(
        l_q=(ctypes.c_double * n)(*rg.l_q),
        bcl_kind=(ctypes.c_int * n)(*[b[0] for b in rg.bcl]),
        bcl_ab=(ctypes.c_double * (2 * n))(*[x for b in rg.bcl for x in b[1:]]),
        bcr_kind=(ctypes.c_int * n)(*[b[0] for b in rg.bcr]),
        bcr_ab=(ctypes.c_double * (2 * n))(*[x for b in rg.bcr for x in b[1:]]),
        bcb_kind=(ctypes.c_int * n)(*rg.bcb),
        bct_kind=(ctypes.c_int * n)(*rg.bct),
        max_wavenumber=(ctypes.c_int * n)(*rg.max_wavenumber),
    )
    d = SsgDesc()
    d.geometry, d.num_cells, d.mubar, d.quadrature = rg.geometry, rg.num_cells, rg.mubar, rg.quadrature
    d.iMin, d.iMax, d.nvars = rg.iMin, rg.iMax, n
    d.kMin, d.kMax, d.kDim, d.b_kDim = rg.kMin, rg.kMax, rg.kDim, rg.b_kDim
    d.ring_kmax, d.patchOffsetL, d.precision = rg.ring_kmax, rg.patchOffsetL, precision
    d.filter_flags = 0   # filter weights go through ssg_plan_set_filter
    d.wavenumber_ahat = 0
    for k, arr in keep.items():
        setattr(d, k, ctypes.cast(arr, _c_int_p if "kind" in k or k == "max_wavenumber" else _c_double_p))
    return d, keep


class Plan:
    """Owns one ssg_plan*."""

    def __init__(self, rg: ResolvedGrid, precision: int = SSG_FP64):
        lib = load()
        self.desc, self._keep = make_desc(rg, precision)
        h = ctypes.c_void_p()
        check(lib.ssg_plan_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.handle = h
        self.precision = precision
        dims = SsgDims()
        check(lib.ssg_plan_get_dims(h, ctypes.byref(dims)))
        self.dims = dims

    def spectral(self, physical_ptr: int, spectral_ptr: int, mem: int, stream: int = 0):
        """ssg_spectral_transform, or ssg_spectral_transform_f32 for an FP32 plan (physical float32)"""
        lib = load()
        fn = lib.ssg_spectral_transform_f32 if self.precision == SSG_FP32_REFINE else lib.ssg_spectral_transform
        check(fn(self.handle, ctypes.c_void_p(physical_ptr), ctypes.c_void_p(spectral_ptr), mem, ctypes.c_void_p(stream)))

    def grid(self, spectral_ptr: int, physical_ptr: int, mem: int, stream: int = 0):
        """ssg_grid_transform, or ssg_grid_transform_f32 for an FP32 plan (physical float32)"""
        lib = load()
        fn = lib.ssg_grid_transform_f32 if self.precision == SSG_FP32_REFINE else lib.ssg_grid_transform
        check(fn(self.handle, ctypes.c_void_p(spectral_ptr), ctypes.c_void_p(physical_ptr), mem, ctypes.c_void_p(stream)))

    def grid_noalloc(self, spectral_ptr: int, physical_ptr: int, stream: int = 0):
        """ssg_grid_transform_noalloc: device pointers, intermediates in the library's static scratch"""
        check(load().ssg_grid_transform_noalloc(self.handle, ctypes.c_void_p(spectral_ptr),
                                                ctypes.c_void_p(physical_ptr), ctypes.c_void_p(stream)))

    def regular(self, spectral_ptr: int, i_pts, j_pts, k_pts, out_ptr: int, mem: int, stream: int = 0):
        """ssg_regular_transform; the point arrays are float64 numpy arrays (or None)."""
        def arr(a):
            if a is None or len(a) == 0:
                return None, 0
            a = np.ascontiguousarray(a, dtype=np.float64)
            return a, len(a)
        (ia, ni), (ja, nj), (ka, nk) = arr(i_pts), arr(j_pts), arr(k_pts)
        ptr = lambda a: a.ctypes.data_as(_c_double_p) if a is not None else None
        check(load().ssg_regular_transform(self.handle, ctypes.c_void_p(spectral_ptr), ptr(ia), ni, ptr(ja), nj,
                                           ptr(ka), nk, ctypes.c_void_p(out_ptr), mem, ctypes.c_void_p(stream)))

    def unstructured(self, spectral_ptr: int, pts: np.ndarray, out_ptr: int, fill: float, oob_error: bool, mem: int,
                     stream: int = 0):
        """ssg_evaluate_unstructured; pts is a float64 (npts, ndim) array (any order)."""
        pts = np.asfortranarray(pts, dtype=np.float64)
        check(load().ssg_evaluate_unstructured(self.handle, ctypes.c_void_p(spectral_ptr),
                                               pts.ctypes.data_as(_c_double_p), pts.shape[0],
                                               ctypes.c_void_p(out_ptr), float(fill), 1 if oob_error else 0, mem,
                                               ctypes.c_void_p(stream)))

    def set_filter(self, fourier_w, chebyshev_w):
        """ssg_plan_set_filter; each table a float64 (nvars, n) array or None."""
        keep = [None if w is None else np.ascontiguousarray(w, dtype=np.float64) for w in (fourier_w, chebyshev_w)]
        ptr = lambda a: a.ctypes.data_as(_c_double_p) if a is not None else None
        check(load().ssg_plan_set_filter(self.handle, ptr(keep[0]), ptr(keep[1])))

    def apply_filter(self, spectral_ptr: int, mem: int, stream: int = 0):
        check(load().ssg_apply_filter(self.handle, ctypes.c_void_p(spectral_ptr), mem, ctypes.c_void_p(stream)))

    def spline_transform(self, shared_ptr: int, spectral_ptr: int, mem: int, stream: int = 0):
        """ssg_spline_transform: B (shared) -> A (spectral)"""
        check(load().ssg_spline_transform(self.handle, ctypes.c_void_p(shared_ptr), ctypes.c_void_p(spectral_ptr),
                                          mem, ctypes.c_void_p(stream)))

    def tile_transform(self, spectral_ptr: int, physical_ptr: int, mem: int, stream: int = 0):
        """ssg_tile_transform: A (spectral) -> physical"""
        check(load().ssg_tile_transform(self.handle, ctypes.c_void_p(spectral_ptr), ctypes.c_void_p(physical_ptr),
                                        mem, ctypes.c_void_p(stream)))

    def set_ahat(self, ahat_ptr, mem: int = SSG_MEM_HOST):
        """ssg_plan_set_ahat; ahat_ptr 0 / None clears."""
        check(load().ssg_plan_set_ahat(self.handle, ctypes.c_void_p(ahat_ptr or None), mem))

    def set_timing(self, on: bool):
        check(load().ssg_plan_set_timing(self.handle, 1 if on else 0))

    def stage_times(self):
        lib = load()
        n = lib.ssg_plan_stage_times(self.handle, None, None, 0)
        names = (ctypes.c_char_p * max(n, 1))()
        ms = (ctypes.c_double * max(n, 1))()
        lib.ssg_plan_stage_times(self.handle, names, ms, n)
        return [(names[i].decode(), ms[i]) for i in range(n)]

    def close(self):
        if getattr(self, "handle", None):
            load().ssg_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owns one ssg_batch*: RL / RLZ grids of one arithmetic form transformed by
    batched launches replayed from a hipGraph.  The grids' device arrays are
    fixed at creation (the tensors are kept alive here)."""

    def __init__(self, grids, physicals, spectrals):
        lib = load()
        n = len(grids)
        self._keep = (list(grids), list(physicals), list(spectrals))
        plans = (ctypes.c_void_p * n)(*[g.plan.handle.value for g in grids])
        ph = (ctypes.c_void_p * n)(*[t.data_ptr() for t in physicals])
        sp = (ctypes.c_void_p * n)(*[t.data_ptr() for t in spectrals])
        h = ctypes.c_void_p()
        check(lib.ssg_batch_create(plans, n, ph, sp, ctypes.byref(h)))
        self.handle = h

    def spectral(self, stream: int = 0):
        check(load().ssg_batch_spectral_transform(self.handle, ctypes.c_void_p(stream)))

    def grid(self, stream: int = 0):
        check(load().ssg_batch_grid_transform(self.handle, ctypes.c_void_p(stream)))

    def tune(self, stream: int = 0):
        check(load().ssg_batch_tune(self.handle, ctypes.c_void_p(stream)))

    def launches(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(load().ssg_batch_launches(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "handle", None):
            load().ssg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_tune(plan: "Plan", physical, spectral, stream: int = 0):
    """Run the launch-shape tuner for one plan on its device arrays (torch tensors)."""
    check(load().ssg_plan_tune(plan.handle, ctypes.c_void_p(physical.data_ptr()),
                               ctypes.c_void_p(spectral.data_ptr()), ctypes.c_void_p(stream)))


TUNING_KEYS = ("syn_ct", "syn_nwq", "ana_mr", "si_seg", "sa_mode")


def plan_tuning(plan: "Plan"):
    """Tuning of a plan (0 = built-in choice): launch shapes; the fifth entry (the spline
    solve) is always the FP64 substitution."""
    out = (ctypes.c_int * 5)()
    check(load().ssg_plan_get_tuning(plan.handle, out))
    return dict(zip(TUNING_KEYS, out))


def plan_set_tuning(plan: "Plan", **choices):
    """Pin tuning choices of a plan by name; the others keep their values."""
    cur = plan_tuning(plan)
    for k, v in choices.items():
        if k not in cur:
            raise KeyError(k)
        cur[k] = int(v)
    arr = (ctypes.c_int * 5)(*[cur[k] for k in TUNING_KEYS])
    check(load().ssg_plan_set_tuning(plan.handle, arr))


def tile_embed(patch: Plan, patch_ptr: int, tile: Plan, tile_ptr: int, spectral_index_l: int, add: bool, mem: int,
               stream: int = 0):
    """ssg_sum_spectral_tile (add) / ssg_set_spectral_tile"""
    lib = load()
    fn = lib.ssg_sum_spectral_tile if add else lib.ssg_set_spectral_tile
    check(fn(patch.handle, ctypes.c_void_p(patch_ptr), tile.handle, ctypes.c_void_p(tile_ptr), int(spectral_index_l),
             mem, ctypes.c_void_p(stream)))


def sum_shared_spectral(shared_ptr: int, shared_rows: int, tile: Plan, tile_ptr: int, spectral_index_l: int,
                        mem: int, stream: int = 0):
    """ssg_sum_shared_spectral"""
    check(load().ssg_sum_shared_spectral(ctypes.c_void_p(shared_ptr), int(shared_rows), tile.handle,
                                         ctypes.c_void_p(tile_ptr), int(spectral_index_l), mem,
                                         ctypes.c_void_p(stream)))
