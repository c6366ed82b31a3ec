"""Pins of the CPU oracle (no GPU needed).

Each oracle function is checked against something that fixes it
independently: the reference's own known-answer tests (test/basis.jl,
test/banded_cholesky.jl, test/mubar.jl, test/transforms.jl), textbook library
routines (numpy.fft, scipy.fft.dct, scipy BSpline, dense SPD solves), closed
forms and invariants (quadrature exactness, polynomial reproduction), and the
independent matrix-form restatement in npref.py for the composed transforms
(which also pins the spectral layouts, TRAP-1 included).
"""
import math

import numpy as np
import pytest
import scipy.fft

import npref
import oracle as O
import springsteel_hip as S
from springsteel_hip.params import compute_derived_params, resolve

RTOL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# --- B-spline basis: test/basis.jl:26-42 known answers ------------------------------------
def test_basis_known_answers():
    xmin, xmax, nc = 0.0, 10.0, 10
    DX = 1.0
    xm = xmin + 3 * DX
    assert abs(O.basis(xmin, xmax, nc, 3, xm, 0) - 4.0 / 6.0) < 1e-14
    assert abs(O.basis(xmin, xmax, nc, 3, xm, 1)) < 1e-14
    assert abs(O.basis(xmin, xmax, nc, 3, xm + 2.0 * DX, 0)) < 1e-14
    assert abs(O.basis(xmin, xmax, nc, 3, xm, 2) - (-2.0)) < 1e-12
    assert math.isnan(O.basis(xmin, xmax, nc, 3, xmin - 0.1, 0))   # DomainError path
    assert math.isnan(O.basis(xmin, xmax, nc, 3, xmax + 0.1, 0))


def test_basis_matches_scipy_bspline():
    sp = npref.NpSpline(0.0, 7.0, 7)
    xs = np.linspace(0.0, 7.0, 301)
    for d in range(4):
        D = sp.design(xs, d)
        for mi in range(sp.M):
            o = np.array([O.basis(0.0, 7.0, 7, mi - 1, x, d) for x in xs])
            if d == 3:   # piecewise constant: compare away from the knots
                mask = np.abs(xs - np.round(xs)) > 1e-9
                assert np.allclose(o[mask], D[mask, mi], atol=1e-12)
            else:
                assert np.allclose(o, D[:, mi], atol=1e-12, rtol=1e-12)


# --- quadrature: test/mubar.jl:42-107 ------------------------------------------------------
@pytest.mark.parametrize("mb", [1, 2, 3, 4, 5])
def test_gauss_quadrature_exactness(mb):
    q, w = O.quadrature(mb, 0)
    assert abs(w.sum() - 1.0) < 1e-15
    for p in range(2 * mb):
        assert abs(np.sum(w * q ** p) - 1.0 / (p + 1)) < 1e-13
    ql, wl = npref.gauss_rule(mb)
    assert np.allclose(q, ql, atol=1e-15) and np.allclose(w, wl, atol=1e-15)


@pytest.mark.parametrize("mb", [1, 2, 4, 7])
def test_regular_quadrature(mb):
    q, w = O.quadrature(mb, 1)
    assert np.allclose(q, (2 * np.arange(1, mb + 1) - 1) / (2 * mb), atol=0)
    assert np.allclose(w, 1.0 / mb, atol=0)


def test_mishpoints_closed_form():
    s = O.OracleSpline(0.0, 20.0, 8)
    x = s.mishpoints()
    DX = 2.5
    ref = np.array([mc * DX + (1 + g) / 2 * DX for mc in range(8) for g in (-math.sqrt(0.6), 0.0, math.sqrt(0.6))])
    assert np.allclose(x, ref, atol=1e-14, rtol=0)


# --- banded Cholesky: test/banded_cholesky.jl:7-42, 81-89 ----------------------------------
def spd_band3(n):
    B = np.zeros((n, n))
    for j in range(1, n + 1):
        for i in range(j, min(n, j + 3) + 1):
            B[i - 1, j - 1] = 1.0 / (1.0 + (i - j)) + 0.1 * (i + j)
    A = B @ B.T
    A = 0.5 * (A + A.T)
    A += 2.0 * np.eye(n)
    return A


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 13, 53, 100, 200])
def test_banded_cholesky_vs_dense(n):
    A = spd_band3(n)
    L = O.band_factor3(A)
    b = np.arange(1, n + 1, dtype=float) + 0.5
    x = O.band_solve3(L, b)
    xr = np.linalg.solve(A, b)
    assert np.allclose(x, xr, rtol=1e-12, atol=1e-12)
    # factor itself against numpy's dense Cholesky
    Ld = np.linalg.cholesky(A)
    for j in range(n):
        for o in range(4):
            if j + o < n:
                assert abs(L[o, j] - Ld[j + o, j]) <= 1e-12 * max(1.0, abs(Ld[j + o, j]))


def test_banded_cholesky_posdef_exception():
    A = spd_band3(6)
    A[3, 3] = -100.0
    with pytest.raises(ArithmeticError):
        O.band_factor3(A)


# --- FFTW definitions ------------------------------------------------------------------------
@pytest.mark.parametrize("n", list(range(1, 41)) + [97, 101, 128, 601, 604, 1204, 2404])
def test_r2hc_matches_numpy_rfft(n):
    x = np.random.default_rng(n).normal(size=n)
    h = O.r2hc(x)
    X = np.fft.rfft(x)
    ref = np.zeros(n)
    ref[: n // 2 + 1] = X.real
    for k in range(1, (n + 1) // 2):
        ref[n - k] = X[k].imag
    assert np.max(np.abs(h - ref)) <= 1e-13 * max(1.0, np.abs(ref).max())
    y = O.hc2r(h)    # unnormalised inverse: HC2R(R2HC(x)) = n x
    assert np.max(np.abs(y - n * x)) <= 1e-12 * n * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("n", list(range(2, 66)))
def test_redft00_matches_scipy_dct1(n):
    x = np.random.default_rng(n).normal(size=n)
    assert np.allclose(O.redft00(x), scipy.fft.dct(x, type=1), rtol=1e-13, atol=1e-13)


# --- spline stages against the independent restatement --------------------------------------
@pytest.mark.parametrize("nc,mb", [(10, 3), (17, 3), (6, 2), (5, 5)])
def test_spline_sb_sa_si_vs_scipy(nc, mb):
    xmin, xmax = 3.0, 3.0 + 2.5 * nc
    s = O.OracleSpline(xmin, xmax, nc, mubar=mb)
    ref = npref.NpSpline(xmin, xmax, nc, mubar=mb)
    x = s.mishpoints()
    assert np.allclose(x, ref.x, atol=1e-13)
    u = np.sin(x / 3.0) + 0.1 * x
    b = s.sb(u)
    assert rel(b, ref.sb(u)) < RTOL
    P, Q = s.pq()
    a = s.sa(b)
    assert rel(a, ref.sa(b)) < 1e-11
    for d in range(3):
        assert rel(s.si(a, x, d), ref.si(a, d)) < 1e-11


def test_spline_alpha1_gamma_fold_vs_dense():
    nc = 12
    s = O.OracleSpline(0.0, 12.0, nc, bcl=(1, -4.0, -1.0), bcr=(1, 0.0, 1.0))
    G = npref.dense_gamma(nc + 3, left=(-4.0, -1.0), right=(0.0, 1.0))
    ref = npref.NpSpline(0.0, 12.0, nc, gamma=G)
    u = np.cos(s.mishpoints() / 2.0)
    b = s.sb(u)
    assert rel(s.sa(b), ref.sa(b)) < 1e-11


def test_quadratic_reproduction_spline():
    """3-point Gauss is exact to degree 5 and Q vanishes on quadratics: SB->SA->SI reproduces
    a quadratic to rounding (the cubic is not reproduced, test/transforms.jl:21-47)."""
    s = O.OracleSpline(0.0, 50.0, 20)
    x = s.mishpoints()
    u = 1.0 + 0.3 * x - 0.01 * x * x
    a = s.sa(s.sb(u))
    assert np.max(np.abs(s.si(a, x, 0) - u)) < 1e-11
    assert np.max(np.abs(s.si(a, x, 1) - (0.3 - 0.02 * x))) < 1e-11
    assert np.max(np.abs(s.si(a, x, 2) + 0.02)) < 1e-11


# --- composed transforms against npref (pins the spectral layouts too) ----------------------
def _grid(gp):
    gp = compute_derived_params(gp)
    rg = resolve(gp)
    return gp, rg, O.OracleGrid(rg)


@pytest.mark.parametrize("nc,off,cap", [(6, 0, -1), (5, 3, -1), (7, 0, 4)])
def test_rl_composed_vs_npref(nc, off, cap):
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RL", num_cells=nc, iMin=0.0, iMax=10.0 * nc,
                                                   spectralIndexL=1 + off // 3, max_wavenumber={"default": cap},
                                                   vars={"u": 1}))
    pts = og.gridpoints()
    rng = np.random.default_rng(nc)
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(pts[:, 0] / 7.0) + pts[:, 0] / 50.0 * (np.cos(pts[:, 1]) + rng.normal() * np.sin(3 * pts[:, 1]))
    spec = og.spectral_transform(phys)
    sp = npref.NpSpline(0.0, 10.0 * nc, nc)
    ri = [r + gp.patchOffsetL for r in range(1, gp.iDim + 1)]
    offs = np.cumsum([0] + [4 + 4 * x for x in ri])
    rings = [phys[offs[i]:offs[i + 1], 0, 0] for i in range(len(ri))]
    sref = npref.rl_forward(sp, ri, rings, [cap] * len(ri))
    assert rel(spec[:, 0], sref) < RTOL
    out = og.grid_transform(spec)
    pref = npref.rl_inverse(sp, ri, spec[:, 0], [cap] * len(ri))
    for s in range(5):
        assert rel(out[:, 0, s], pref[s]) < 1e-11, s


def test_rz_composed_vs_npref():
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RZ", num_cells=6, iMin=0.0, iMax=30.0, kMin=0.0,
                                                   kMax=15.0, kDim=12, vars={"u": 1}))
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = np.sin(pts[:, 0] / 9.0) * np.cos(pts[:, 1] / 5.0) + 0.01 * pts[:, 1] ** 2
    spec = og.spectral_transform(phys)
    sp = npref.NpSpline(0.0, 30.0, 6)
    ch = npref.NpCheb(0.0, 15.0, 12, gp.b_kDim)
    assert rel(spec[:, 0], npref.rz_forward(sp, ch, phys[:, 0, 0])) < RTOL
    out = og.grid_transform(spec)
    pref = npref.rz_inverse(sp, ch, spec[:, 0])
    for s in range(5):
        assert rel(out[:, 0, s], pref[s]) < 1e-11, s


@pytest.mark.parametrize("off", [0, 3])
def test_rlz_composed_vs_npref(off):
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=60.0, kMin=0.0,
                                                   kMax=10.0, kDim=8, spectralIndexL=1 + off // 3,
                                                   vars={"u": 1}))
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = (np.sin(pts[:, 0] / 20.0) + pts[:, 0] / 60.0 * np.cos(2 * pts[:, 1])) * (1 + pts[:, 2] / 10.0) \
        + 0.02 * pts[:, 2] ** 2
    spec = og.spectral_transform(phys)
    sp = npref.NpSpline(0.0, 60.0, 4)
    ch = npref.NpCheb(0.0, 10.0, 8, gp.b_kDim)
    ri = [r + gp.patchOffsetL for r in range(1, gp.iDim + 1)]
    assert rel(spec[:, 0], npref.rlz_forward(sp, ch, ri, phys[:, 0, 0])) < RTOL
    out = og.grid_transform(spec)
    pref = npref.rlz_inverse(sp, ch, ri, spec[:, 0])
    for s in range(7):
        assert rel(out[:, 0, s], pref[s]) < 1e-11, s


# --- reference integration tolerances (test/transforms.jl) ----------------------------------
def test_reference_rl_roundtrips():
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1}))
    pts = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = pts[:, 0]
    out = og.grid_transform(og.spectral_transform(phys))
    assert np.max(np.abs(out[:, 0, 0] - pts[:, 0])) < 1e-3
    assert np.max(np.abs(out[:, 0, 1] - 1.0)) < 0.05
    assert np.max(np.abs(out[:, 0, 2])) < 0.05
    phys[:, 0, 0] = pts[:, 0] * np.cos(pts[:, 1])
    out = og.grid_transform(og.spectral_transform(phys))
    assert np.max(np.abs(out[:, 0, 0] - phys[:, 0, 0])) < 1e-2


def test_polynomial_reproduction_rlz():
    """Band-limited in lambda, quadratic in r, degree < b_kDim in z: reproduced to rounding
    with exact derivatives in every slot."""
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RLZ", num_cells=6, iMin=0.0, iMax=30.0, kMin=0.0,
                                                   kMax=10.0, kDim=10, vars={"u": 1}))
    p = og.gridpoints()
    r, lam, z = p[:, 0], p[:, 1], p[:, 2]
    # wavenumber 1 is carried by every ring (ring r holds k <= r)
    f = (r * r) * np.cos(lam) * (1 + z / 10.0) ** 2
    phys = og.new_physical()
    phys[:, 0, 0] = f
    out = og.grid_transform(og.spectral_transform(phys))
    zf = (1 + z / 10.0) ** 2
    expect = [f, 2 * r * np.cos(lam) * zf, 2 * np.cos(lam) * zf, -r * r * np.sin(lam) * zf,
              -f, r * r * np.cos(lam) * 2 * (1 + z / 10.0) / 10.0, r * r * np.cos(lam) * 2 / 100.0]
    for s in range(7):
        assert rel(out[:, 0, s], expect[s]) < 1e-10, s


def test_noise_floor_between_cpu_restatements():
    """On a field linear in z, d2/dz2 is exactly zero and any FP64 implementation returns
    rounding noise.  The oracle and the independent numpy restatement differ there by the
    same order as the GPU does (measured against the slot's scale floor), which fixes the
    tolerance used for numerically-zero slots in the GPU parity tests."""
    from parity_util import TOL_NOISE_FLOOR, zero_truth_slots
    from springsteel_hip.metrics import rel_l2_slots
    gp, rg, og = _grid(S.SpringsteelGridParameters(geometry="RLZ", num_cells=6, iMin=0.0, iMax=60.0, kMin=0.0,
                                                   kMax=20000.0, kDim=32, vars={"u": 1}))
    p = og.gridpoints()
    phys = og.new_physical()
    phys[:, 0, 0] = (2 * (p[:, 0] / 50) / (1 + (p[:, 0] / 50) ** 2) + 0.05 * np.cos(3 * p[:, 1])) \
        * (1 - p[:, 2] / 40000.0)
    spec = og.spectral_transform(phys)
    out = og.grid_transform(spec)
    sp = npref.NpSpline(0.0, 60.0, 6)
    ch = npref.NpCheb(0.0, 20000.0, 32, gp.b_kDim)
    ri = list(range(1, gp.iDim + 1))
    # each restatement runs its own forward transform, as the GPU path does
    nspec = npref.rlz_forward(sp, ch, ri, phys[:, 0, 0])
    pref = np.stack(npref.rlz_inverse(sp, ch, ri, nspec), axis=1)
    assert (0, 6) in zero_truth_slots(out, "RLZ", gp)
    e = rel_l2_slots(out[:, 0, :], pref, "RLZ", gp)
    print("oracle vs numpy restatement, floored rel L2 per slot:", ["%.1e" % x for x in e])
    assert max(e[:6]) < 1e-11
    assert e[6] < TOL_NOISE_FLOOR
