"""ctypes wrapper of the CPU oracle (libssg_oracle.so, built from ssg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker / the timed CPU port.

The oracle takes the per-variable resolved grid description (any object with
the attributes of springsteel_hip.params.ResolvedGrid) and runs the literal
restatement of the reference's loops on host arrays in Julia column-major
layout.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libssg_oracle.so")

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class _Params(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("num_cells", ctypes.c_int), ("mubar", ctypes.c_int),
                ("quadrature", ctypes.c_int), ("iMin", ctypes.c_double), ("iMax", ctypes.c_double),
                ("nvars", ctypes.c_int), ("l_q", _dp), ("bcl_kind", _ip), ("bcl_ab", _dp), ("bcr_kind", _ip),
                ("bcr_ab", _dp), ("kMin", ctypes.c_double), ("kMax", ctypes.c_double), ("kDim", ctypes.c_int),
                ("b_kDim", ctypes.c_int), ("bcb_kind", _ip), ("bct_kind", _ip), ("max_wavenumber", _ip),
                ("ring_kmax", ctypes.c_int), ("patchOffsetL", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.ssgo_grid_create.restype = ctypes.c_void_p
        L.ssgo_grid_create.argtypes = [ctypes.POINTER(_Params)]
        L.ssgo_grid_destroy.argtypes = [ctypes.c_void_p]
        L.ssgo_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.ssgo_error.restype = ctypes.c_char_p
        L.ssgo_spectral_transform.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ssgo_grid_transform.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ssgo_spline_transform.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ssgo_tile_transform.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.ssgo_gridpoints.argtypes = [ctypes.c_void_p, _dp]
        L.ssgo_regular_transform.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                             ctypes.c_int, _dp, ctypes.c_int]
        L.ssgo_set_dzz_bound.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ssgo_set_ahat.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ssgo_spline_sa_ahat.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ssgo_evaluate_unstructured.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_long, _dp, ctypes.c_double,
                                                 ctypes.c_int]
        L.ssgo_cheb_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int]
        L.ssgo_cheb_eval.restype = ctypes.c_double
        L.ssgo_basis.restype = ctypes.c_double
        L.ssgo_basis.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int]
        L.ssgo_quadrature.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        for f in (L.ssgo_r2hc, L.ssgo_hc2r, L.ssgo_redft00):
            f.argtypes = [ctypes.c_int, _dp, _dp]
        L.ssgo_band_factor3.argtypes = [ctypes.c_int, _dp, _dp]
        L.ssgo_band_solve3.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.ssgo_spline_create.restype = ctypes.c_void_p
        L.ssgo_spline_create.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.ssgo_spline_destroy.argtypes = [ctypes.c_void_p]
        L.ssgo_spline_mdim.argtypes = [ctypes.c_void_p]
        L.ssgo_spline_minterior.argtypes = [ctypes.c_void_p]
        L.ssgo_spline_mishpoints.argtypes = [ctypes.c_void_p, _dp]
        L.ssgo_spline_pq.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.ssgo_spline_band.argtypes = [ctypes.c_void_p, _dp]
        L.ssgo_spline_sb.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.ssgo_spline_sa.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.ssgo_spline_si.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class OracleGrid:
    """One grid of the CPU restatement."""

    def __init__(self, rg):
        n = rg.nvars
        self._keep = [
            np.asarray(rg.l_q, dtype=np.float64),
            np.asarray([b[0] for b in rg.bcl], dtype=np.int32),
            np.asarray([x for b in rg.bcl for x in b[1:]], dtype=np.float64),
            np.asarray([b[0] for b in rg.bcr], dtype=np.int32),
            np.asarray([x for b in rg.bcr for x in b[1:]], dtype=np.float64),
            np.asarray(rg.bcb, dtype=np.int32),
            np.asarray(rg.bct, dtype=np.int32),
            np.asarray(rg.max_wavenumber, dtype=np.int32),
        ]
        k = self._keep
        p = _Params(rg.geometry, rg.num_cells, rg.mubar, rg.quadrature, rg.iMin, rg.iMax, n,
                    k[0].ctypes.data_as(_dp), k[1].ctypes.data_as(_ip), k[2].ctypes.data_as(_dp),
                    k[3].ctypes.data_as(_ip), k[4].ctypes.data_as(_dp), rg.kMin, rg.kMax, rg.kDim, rg.b_kDim,
                    k[5].ctypes.data_as(_ip), k[6].ctypes.data_as(_ip), k[7].ctypes.data_as(_ip),
                    rg.ring_kmax, rg.patchOffsetL)
        self.h = lib().ssgo_grid_create(ctypes.byref(p))
        if not self.h:
            raise ValueError(lib().ssgo_error().decode())
        d = (ctypes.c_long * 8)()
        lib().ssgo_dims(self.h, d)
        self.npts, self.nspec, self.nslots = d[0], d[1], d[2]
        self.iDim, self.b_iDim, self.jDim, self.kDim, self.b_kDim = d[3], d[4], d[5], d[6], d[7]
        self.nvars = n
        self.geometry = rg.geometry
        self.patchOffsetL = rg.patchOffsetL

    def apply_filter(self, spectral, fourier_w=None, chebyshev_w=None):
        """_applyFilter_impl! (src/filtering.jl:290-428) on spectral (nspec, nvars) in place,
        given the per-variable weights _filter_weight would give: fourier_w[v][k],
        k = 0..iDim + patchOffsetL, and chebyshev_w[v][z_b - 1]; None for an empty dict
        (1.0 rows for the variables a dict does not cover).  Ranges are the reference's
        1-based inclusive block bounds."""
        b, K = self.b_iDim, self.iDim + self.patchOffsetL
        rtol = math.sqrt(np.finfo(np.float64).eps)

        def block(r1, r2, v, w):
            if w == 1.0 or (math.isfinite(w) and abs(w - 1.0) <= rtol * max(abs(w), 1.0)):   # w ≈ 1.0
                return
            if w == 0.0:                                                                      # w ≈ 0.0
                spectral[r1 - 1:r2, v] = 0.0
            else:
                spectral[r1 - 1:r2, v] *= w

        g = self.geometry
        if g == 2:                                     # RZ: 290-314
            if chebyshev_w is None:
                return spectral
            for v in range(self.nvars):
                for z in range(1, self.b_kDim + 1):
                    block((z - 1) * b + 1, z * b, v, float(chebyshev_w[v][z - 1]))
        elif g in (1, 4):                              # RL / SL: 320-364
            if fourier_w is None:
                return spectral
            for v in range(self.nvars):
                block(1, b, v, float(fourier_w[v][0]))
                for k in range(1, K + 1):
                    w = float(fourier_w[v][k])
                    block((2 * k - 1) * b + 1, 2 * k * b, v, w)
                    block(2 * k * b + 1, (2 * k + 1) * b, v, w)
        else:                                          # RLZ / SLZ: 372-428
            if fourier_w is None and chebyshev_w is None:
                return spectral
            block_size = b * (1 + K * 2)
            for v in range(self.nvars):
                for z_b in range(1, self.b_kDim + 1):
                    c_w = 1.0 if chebyshev_w is None else float(chebyshev_w[v][z_b - 1])
                    block_start = (z_b - 1) * block_size
                    f_w0 = 1.0 if fourier_w is None else float(fourier_w[v][0])
                    block(block_start + 1, block_start + b, v, c_w * f_w0)
                    for k in range(1, K + 1):
                        f_wk = 1.0 if fourier_w is None else float(fourier_w[v][k])
                        w = c_w * f_wk
                        r1_real = block_start + b + (k - 1) * 2 * b + 1
                        r2_real = r1_real + b - 1
                        block(r1_real, r2_real, v, w)
                        block(r2_real + 1, r2_real + b, v, w)
        return spectral

    def new_physical(self):
        return np.zeros((self.npts, self.nvars, self.nslots), order="F")

    def new_spectral(self):
        return np.zeros((self.nspec, self.nvars), order="F")

    def spectral_transform(self, physical, spectral=None, threads=1):
        if spectral is None:
            spectral = self.new_spectral()
        assert physical.flags.f_contiguous and spectral.flags.f_contiguous
        lib().ssgo_spectral_transform(self.h, _p(physical), _p(spectral), threads)
        return spectral

    def grid_transform(self, spectral, physical=None, threads=1):
        if physical is None:
            physical = self.new_physical()
        assert physical.flags.f_contiguous and spectral.flags.f_contiguous
        lib().ssgo_grid_transform(self.h, _p(spectral), _p(physical), threads)
        return physical

    def spline_transform(self, shared, spectral=None, threads=1):
        """splineTransform!: SA of every line of the B coefficients `shared` -> A coefficients."""
        if spectral is None:
            spectral = self.new_spectral()
        assert shared.flags.f_contiguous and spectral.flags.f_contiguous
        if lib().ssgo_spline_transform(self.h, _p(shared), _p(spectral), threads) != 0:
            raise RuntimeError(lib().ssgo_error().decode())
        return spectral

    def tile_transform(self, spectral, physical=None, threads=1):
        """tileTransform!: every physical slot from the A coefficients `spectral` (no SA)."""
        if physical is None:
            physical = self.new_physical()
        assert physical.flags.f_contiguous and spectral.flags.f_contiguous
        if lib().ssgo_tile_transform(self.h, _p(spectral), _p(physical), threads) != 0:
            raise RuntimeError(lib().ssgo_error().decode())
        return physical

    def regular_transform(self, spectral, i_pts, j_pts=None, k_pts=None, threads=1):
        """regularGridTransform at the tensor product of the points: (npr, nvars, nslots), Fortran order."""
        assert spectral.flags.f_contiguous
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else np.zeros(1) for a in
                (i_pts, j_pts, k_pts)]
        ns = [len(a) if a is not None else 0 for a in (i_pts, j_pts, k_pts)]
        npr = int(np.prod([n for n in ns if n > 0]))
        out = np.zeros((npr, self.nvars, int(self.nslots)), order="F")
        lib().ssgo_regular_transform(self.h, _p(spectral), _p(arrs[0]), ns[0], _p(arrs[1]), ns[1], _p(arrs[2]),
                                     ns[2], _p(out), threads)
        return out

    def set_ahat(self, ahat):
        """R3X background coefficients in the layout of spectral (nspec, nvars), or None."""
        if ahat is None:
            lib().ssgo_set_ahat(self.h, None)
            return
        ahat = np.asfortranarray(ahat, dtype=np.float64)
        assert ahat.shape == (self.nspec, self.nvars)
        lib().ssgo_set_ahat(self.h, ahat.ctypes.data)

    def evaluate_unstructured(self, spectral, pts, fill=float("nan"), threads=1):
        """evaluate_unstructured: values (n, nvars), Fortran order, at the rows of pts (n, ndim)."""
        assert spectral.flags.f_contiguous
        pts = np.asfortranarray(pts, dtype=np.float64)
        out = np.zeros((pts.shape[0], self.nvars), order="F")
        lib().ssgo_evaluate_unstructured(self.h, _p(spectral), _p(pts), pts.shape[0], _p(out), float(fill), threads)
        return out

    def regular_dzz_bound(self, spectral, i_pts, j_pts=None, k_pts=None, threads=1):
        """The scale of the rounding error of the reference's FP64 d2/dz2 formula at each
        regular point (eps times it bounds the error of its final difference), in the
        d2/dz2 slot of an array laid out as regular_transform's."""
        lib().ssgo_set_dzz_bound(self.h, 1)
        try:
            return self.regular_transform(spectral, i_pts, j_pts, k_pts, threads)
        finally:
            lib().ssgo_set_dzz_bound(self.h, 0)

    def cheb_eval(self, v, b, z, derivative=0, dzz_bound=False):
        """One Chebyshev column b of variable v: CAtransform! then the value (or d/dz, d2/dz2) at z.
        dzz_bound with derivative 2: the rounding scale of regular_dzz_bound instead."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().ssgo_set_dzz_bound(self.h, 1 if dzz_bound else 0)
        try:
            return lib().ssgo_cheb_eval(self.h, v, _p(b), float(z), derivative)
        finally:
            lib().ssgo_set_dzz_bound(self.h, 0)

    def gridpoints(self):
        ndim = 3 if self.geometry in (3, 5) else 2
        pts = np.zeros((self.npts, ndim), order="F")
        lib().ssgo_gridpoints(self.h, _p(pts))
        return pts

    def __del__(self):
        try:
            if self.h:
                lib().ssgo_grid_destroy(self.h)
                self.h = None
        except Exception:
            pass


class OracleSpline:
    """One Spline1D of the restatement (component-level pinning)."""

    def __init__(self, xmin, xmax, num_cells, mubar=3, quadrature=0, l_q=2.0, bcl=(0, 0.0, 0.0), bcr=(0, 0.0, 0.0)):
        self.h = lib().ssgo_spline_create(xmin, xmax, num_cells, mubar, quadrature, l_q, bcl[0], bcl[1], bcl[2],
                                          bcr[0], bcr[1], bcr[2])
        if not self.h:
            raise ValueError(lib().ssgo_error().decode())
        self.Mdim = lib().ssgo_spline_mdim(self.h)
        self.Min = lib().ssgo_spline_minterior(self.h)
        self.mishDim = num_cells * mubar

    def mishpoints(self):
        x = np.zeros(self.mishDim)
        lib().ssgo_spline_mishpoints(self.h, _p(x))
        return x

    def pq(self):
        P = np.zeros((self.Mdim, self.Mdim), order="F")
        Q = np.zeros((self.Mdim, self.Mdim), order="F")
        lib().ssgo_spline_pq(self.h, _p(P), _p(Q))
        return P, Q

    def band(self):
        L = np.zeros((4, self.Min), order="F")
        lib().ssgo_spline_band(self.h, _p(L))
        return L

    def sb(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        b = np.zeros(self.Mdim)
        lib().ssgo_spline_sb(self.h, _p(u), _p(b))
        return b

    def sa(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        a = np.zeros(self.Mdim)
        lib().ssgo_spline_sa(self.h, _p(b), _p(a))
        return a

    def sa_ahat(self, b, ahat):
        """SAtransform! with the R3X background coefficients ahat (used on an R3X spline)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        ahat = np.ascontiguousarray(ahat, dtype=np.float64)
        a = np.zeros(self.Mdim)
        lib().ssgo_spline_sa_ahat(self.h, _p(b), _p(ahat), _p(a))
        return a

    def si(self, a, pts, derivative=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        u = np.zeros(len(pts))
        lib().ssgo_spline_si(self.h, _p(a), _p(pts), len(pts), derivative, _p(u))
        return u

    def __del__(self):
        try:
            if self.h:
                lib().ssgo_spline_destroy(self.h)
                self.h = None
        except Exception:
            pass


def basis(xmin, xmax, num_cells, m, x, derivative):
    return lib().ssgo_basis(xmin, xmax, num_cells, m, x, derivative)


def quadrature(mubar, quad=0):
    q = np.zeros(mubar)
    w = np.zeros(mubar)
    lib().ssgo_quadrature(mubar, quad, _p(q), _p(w))
    return q, w


def r2hc(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().ssgo_r2hc(len(x), _p(x), _p(out))
    return out


def hc2r(h):
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros_like(h)
    lib().ssgo_hc2r(len(h), _p(h), _p(out))
    return out


def redft00(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().ssgo_redft00(len(x), _p(x), _p(out))
    return out


def band_factor3(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((4, n), order="F")
    rc = lib().ssgo_band_factor3(n, _p(A), _p(L))
    if rc != 0:
        raise ArithmeticError(f"PosDefException({-rc})")
    return L


def band_solve3(L, b):
    L = np.asfortranarray(L, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().ssgo_band_solve3(L.shape[1], _p(L), _p(b), _p(x))
    return x


# ---------------------------------------------------------------------------
# The tile sums of src/tiling.jl on numpy arrays (column-major (nspec, nvars)
# spectral arrays, 1-based indices as in the reference, shifted by one where
# they index).  A grid is described by (geometry, b_iDim, iDim + patchOffsetL,
# b_kDim); geometry "RL" / "SL" or "RLZ" / "SLZ".
# ---------------------------------------------------------------------------
def _tile_blocks(patch_dims, tile_dims, siL):
    """(patch row, tile row, length) of every block that sumSpectralTile! /
    setSpectralTile! move: src/tiling.jl:1048-1071 (RL), 1074-1098 (SL),
    1101-1134 (RLZ), 1137-1167 (SLZ); the set forms 1204-1325 walk the same."""
    geo, bp, Kp, _ = patch_dims
    _, bt, Kt, bkt = tile_dims
    out = []
    if geo in ("RL", "SL"):
        siR = siL + bt - 1
        out.append((siL, 1, siR - siL + 1))                  # wavenumber 0
        for k in range(1, Kt + 1):
            p = k * 2
            t = k * 2
            out.append(((p - 1) * bp + siL, (t - 1) * bt + 1, bt))   # real
            out.append((p * bp + siL, t * bt + 1, bt))               # imaginary
    else:
        wn_p = bp * (1 + 2 * Kp)
        wn_t = bt * (1 + 2 * Kt)
        for z_b in range(1, bkt + 1):
            zp = (z_b - 1) * wn_p
            zt = (z_b - 1) * wn_t
            out.append((zp + siL, zt + 1, bt))
            for k in range(1, Kt + 1):
                p = (k - 1) * 2
                out.append((zp + (p + 1) * bp + siL, zt + (p + 1) * bt + 1, bt))
                out.append((zp + (p + 2) * bp + siL, zt + (p + 2) * bt + 1, bt))
    return out


def sum_spectral_tile(patch_spec, patch_dims, tile_spec, tile_dims, siL):
    """sumSpectralTile!(patch, tile) in place on patch_spec."""
    for pr, tr, n in _tile_blocks(patch_dims, tile_dims, siL):
        patch_spec[pr - 1:pr - 1 + n, :] += tile_spec[tr - 1:tr - 1 + n, :]
    return patch_spec


def set_spectral_tile(patch_spec, patch_dims, tile_spec, tile_dims, siL):
    """setSpectralTile!(patch, tile) in place on patch_spec."""
    patch_spec[:] = 0.0
    for pr, tr, n in _tile_blocks(patch_dims, tile_dims, siL):
        patch_spec[pr - 1:pr - 1 + n, :] = tile_spec[tr - 1:tr - 1 + n, :]
    return patch_spec


def sum_shared_spectral(shared, tile_spec, tile_b_iDim, siL):
    """sumSharedSpectral's generic method (src/tiling.jl:1379-1389) in place on shared."""
    siR = siL + tile_b_iDim - 1 - 3
    if siR >= siL:
        shared[siL - 1:siR, :] = tile_spec[0:tile_b_iDim - 3, :]
    return shared
