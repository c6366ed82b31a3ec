"""Pins of the oracle's boundary-condition folds from what each fold enforces
(no GPU needed).

The reference's own tests check only that these folds give finite numbers
(test/basis.jl:288-304, 1208-1228), so they are pinned here by the
mathematics of each fold, through the oracle's public entry points:

* Cubic B-spline Gamma (src/CubicBSpline.jl:321-500, calcGammaBC 648-706).
  With the basis values (1, 4, 1)/6, slopes (-1, 0, 1)/(2 DX) and curvatures
  (1, -2, 1)/DX^2 at a knot, the rank-1 folds R1T0 / R1T1 / R1T2 make the
  value / slope / curvature of the spline vanish at that end, and the rank-2
  folds R2T10 / R2T20 make value and slope / value and curvature vanish.  So
  after SA on a random right-hand side, SI at the boundary must give zero in
  exactly those orders (and not in the others, which shows the check bites).

* Chebyshev gammaBC (src/Chebyshev.jl:991-1177), applied by CAtransform
  (473-499) as a = bfill + gamma' bfill.  With the DCT-I synthesis
  u_j = sum_k w_k a_k cos(pi j k / (N-1)), w = (1, 2, ..., 2, 1), the bottom
  (j = 0, x = +1) and top (x = -1) of the column:
    R1T0 bottom: a - bfill is a constant vector and u(bottom) = 0;
    R1T0 top:    a - bfill is proportional to (-1)^k and u(top) = 0;
    R1T1 bottom (Wang et al. 1993 global coefficient method): a - bfill is
                 proportional to 1/c_k (c = (2, 1, ..., 1, 2)) and
                 sum_k k^2 a_k = 0 (the slope at x = +1 of sum_k a_k T_k);
    R1T1 top:    a - bfill is proportional to (-1)^k / c_k and
                 sum_k (-1)^(k+1) k^2 a_k = 0 (the slope at x = -1).
  Each pair of properties fixes its rank-1 fold uniquely.  The four mixed
  combinations are built in the reference as the sum of the two single
  matrices, so their correction is the sum of the two single corrections.
  The coefficients a are recovered from the oracle's own evaluation of the
  folded column at the N Chebyshev-Gauss-Lobatto points (u = DCT-I(a), and
  DCT-I is its own inverse up to 2 (N - 1)).
"""
import numpy as np
import pytest
import scipy.fft

import oracle as O
import springsteel_hip as S
from springsteel_hip.params import compute_derived_params, resolve

CB = S.CubicBSpline
CH = S.Chebyshev

# spline BC kinds of the oracle (ssg_oracle.h): 1 = alpha1 (rank 1), 2 = alpha2 (rank 2)
SPLINE_BCS = {
    "R1T0": (1, -4.0, -1.0, {0}),
    "R1T1": (1, 0.0, 1.0, {1}),
    "R1T2": (1, 2.0, -1.0, {2}),
    "R2T10": (2, 1.0, -0.5, {0, 1}),
    "R2T20": (2, -1.0, 0.0, {0, 2}),
}


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("name", sorted(SPLINE_BCS))
def test_spline_fold_enforces_its_conditions(name, side):
    kind, a1, b1, zero_orders = SPLINE_BCS[name]
    nc, xmin, xmax = 17, 3.0, 37.0
    bc = (kind, a1, b1)
    s = O.OracleSpline(xmin, xmax, nc, bcl=bc if side == "left" else (0, 0.0, 0.0),
                       bcr=bc if side == "right" else (0, 0.0, 0.0))
    dx = (xmax - xmin) / nc
    rng = np.random.default_rng(3 + kind)
    x = s.mishpoints()
    for trial in range(4):
        u = rng.normal(size=x.size) + np.cos(x / 5.0)
        a = s.sa(s.sb(u))
        end = np.array([xmin if side == "left" else xmax])
        scale = np.abs(a).max()
        for d in range(3):
            val = s.si(a, end, d)[0]
            unit = scale / dx ** d            # size of one basis term of order d at the knot
            if d in zero_orders:
                assert abs(val) <= 1e-12 * unit, (name, side, d, val, unit)
            else:
                # the orders the fold leaves free are not zero: the check above bites
                assert abs(val) > 1e-6 * unit, (name, side, d, val, unit)


def _cheb_grid(combos, kDim=12):
    names = {f"v{i}": i + 1 for i in range(len(combos))}
    bcb = {f"v{i}": c[0] for i, c in enumerate(combos)}
    bct = {f"v{i}": c[1] for i, c in enumerate(combos)}
    gp = S.SpringsteelGridParameters(geometry="RZ", num_cells=4, iMin=0.0, iMax=4.0, kMin=-3.0, kMax=7.0,
                                     kDim=kDim, vars=names, BCB=bcb, BCT=bct)
    rg = resolve(compute_derived_params(gp))
    return O.OracleGrid(rg), rg


def _folded_coefficients(og, rg, v, b):
    """a of CAtransform from the oracle's values at the CGL points (ascending z, so the
    first point is the bottom, x = +1)."""
    N = rg.kDim
    zp = np.cos(np.arange(N) * np.pi / (N - 1)) * (-0.5 * (rg.kMax - rg.kMin)) + 0.5 * (rg.kMin + rg.kMax)
    u = np.array([og.cheb_eval(v, b, z, 0) for z in zp])
    return scipy.fft.dct(u, type=1) / (2 * (N - 1)), u


COMBOS = [(CH.R0, CH.R0), (CH.R1T0, CH.R0), (CH.R0, CH.R1T0), (CH.R1T1, CH.R0), (CH.R0, CH.R1T1),
          (CH.R1T0, CH.R1T0), (CH.R1T1, CH.R1T1), (CH.R1T0, CH.R1T1), (CH.R1T1, CH.R1T0)]


@pytest.mark.parametrize("kDim", [9, 12, 25])
def test_chebyshev_single_folds(kDim):
    og, rg = _cheb_grid(COMBOS, kDim)
    N, bk = rg.kDim, rg.b_kDim
    k = np.arange(N)
    c = np.ones(N)
    c[0] = c[-1] = 2.0
    alt = (-1.0) ** k
    rng = np.random.default_rng(kDim)
    for trial in range(3):
        b = rng.normal(size=bk)
        bfill = np.concatenate([b, np.zeros(N - bk)])
        scale = np.abs(bfill).sum()
        # the oracle evaluates the two end points at an angle clamped 1e-14 inside [-1, 1]
        # (src/transforms_cylindrical.jl _cheb_eval_pts!), which moves those values by up
        # to their slope (~N^2 scale) times that margin; the recovered a inherits it
        tol = 1e-14 * scale * N * N
        for v, (bot, top) in enumerate(COMBOS[:5]):
            a, u = _folded_coefficients(og, rg, v, b)
            corr = a - bfill
            if (bot, top) == (CH.R0, CH.R0):
                assert np.abs(corr).max() <= tol
                continue
            if bot == CH.R1T0:
                direction, zero = np.ones(N), abs(u[0])
            elif top == CH.R1T0:
                direction, zero = alt, abs(u[-1])
            elif bot == CH.R1T1:
                direction, zero = 1.0 / c, abs(np.dot(k * k, a))
            else:
                direction, zero = alt / c, abs(np.dot(-alt * k * k, a))
            # the correction lies along the fold's direction ...
            t = np.dot(corr, direction) / np.dot(direction, direction)
            assert np.abs(corr - t * direction).max() <= tol, (v, bot, top)
            assert abs(t) > 1e-6 * scale                        # and is not empty
            # ... and the fold's condition holds (the slope functionals weigh a_k by k^2)
            assert zero <= 10 * tol * (N * N if bot == CH.R1T1 or top == CH.R1T1 else 1), (v, bot, top, zero)


def test_chebyshev_combined_folds_add_the_single_corrections():
    og, rg = _cheb_grid(COMBOS, 14)
    N, bk = rg.kDim, rg.b_kDim
    key = [(str(bot), str(top)) for bot, top in COMBOS]
    r0 = str(CH.R0)
    rng = np.random.default_rng(5)
    for trial in range(3):
        b = rng.normal(size=bk)
        bfill = np.concatenate([b, np.zeros(N - bk)])
        corr = {key[v]: _folded_coefficients(og, rg, v, b)[0] - bfill for v in range(len(COMBOS))}
        for bot, top in key[5:]:
            want = corr[(bot, r0)] + corr[(r0, top)]
            got = corr[(bot, top)]
            assert np.abs(got - want).max() <= 3e-14 * np.abs(bfill).sum() * N * N, (bot, top)
