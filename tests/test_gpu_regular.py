"""GPU: regularGridTransform (ssg_regular_transform) against the oracle.

Bar: relative L2 per (variable, slot) <= 1e-12, with the scale floors of
springsteel_hip.metrics for derivative slots (a slot whose exact value is zero
holds only rounding).  The oracle's restatement is pinned in
test_regular_pins.py, including the reference's mirrored sine convention,
which the library reproduces.

One exception, from the reference's own formula: at a z whose Chebyshev angle
is clamped (|xi| >= 1 - 1e-10, the domain ends), _cheb_dzz_pts! divides a
difference of two O(t) terms by sin(t)^3 with t ~ 1.4e-5, so the rounding of
that difference is far above eps times the value (up to 1e-6 relative here).
The library evaluates the same FP64 expressions (so the larger rounding, from
sin / cos of the FP64 product k t, is the reference's own), but sums the
series in another order; the d2/dz2 slot at those z is held to 4 eps times the
rounding scale of that final difference, as the oracle reports it
(regular_dzz_bound), plus 1e-12 of the slot's norm for the upstream stages.
Every other entry, and d2/dz2 at every other z, is held to 1e-12.
"""
import math

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.errors import DomainError
from springsteel_hip.metrics import rel_l2_slots
from parity_util import seeded_field

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _run(gp, seed=0, points=None):
    g = S.createGrid(gp)
    og = O.OracleGrid(g.resolved)
    pts = S.getGridpoints(g)
    nv = g.resolved.nvars
    g.physical[:, :, 0] = seeded_field(pts, nv, seed, g.geometry)
    S.spectralTransform_(g)
    spec = np.asfortranarray(g.spectral.copy())
    axes = points if points is not None else S.grid._regular_points(g, (S.getRegularGridpoints(g),))
    out = S.regularGridTransform(g, *[a for a in axes if a is not None])
    ref = og.regular_transform(spec, *axes, threads=8)
    g._oracle = (og, spec)
    return g, out, ref, axes


def _clamped_rows(g, axes):
    """Rows whose z has a clamped Chebyshev angle (the d2/dz2 exception above)."""
    i, j, k = axes
    if k is None:
        return None
    p = g.params
    scale, offset = -0.5 * (p.kMax - p.kMin), 0.5 * (p.kMin + p.kMax)
    zc = np.abs((np.asarray(k) - offset) / scale) >= 1.0 - 1e-10
    reps = len(i) * (len(j) if j is not None else 1)
    return np.tile(zc, reps)


def _check(g, out, ref, axes):
    p = g.params
    worst = 0.0
    clamp = _clamped_rows(g, axes)
    bound = None
    if clamp is not None and clamp.any():
        og, spec = g._oracle
        bound = og.regular_dzz_bound(spec, *axes, threads=8)
    for v in range(g.resolved.nvars):
        if bound is not None:
            s2 = out.shape[2] - 1                         # d2/dz2 slot
            o, r = out[:, v, :].copy(), ref[:, v, :].copy()
            err = np.abs(o[clamp, s2] - r[clamp, s2])
            lim = 4.0 * np.finfo(np.float64).eps * bound[clamp, v, s2] + 1e-12 * np.linalg.norm(r[:, s2])
            assert np.all(err <= lim), (v, float(np.max(err / lim)))
            o[clamp, s2] = r[clamp, s2]
        else:
            o, r = out[:, v, :], ref[:, v, :]
        errs = rel_l2_slots(o, r, g.geometry, p)
        worst = max(worst, max(errs))
        assert max(errs) <= TOL, (v, errs)
    return worst


def test_rl_regular_default_points():
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=12, iMin=0.0, iMax=60.0,
                                     vars={"u": 1, "v": 2, "w": 3}, BCL={"default": S.CubicBSpline.R0,
                                                                          "v": S.CubicBSpline.R1T0},
                                     BCR={"default": S.CubicBSpline.R0, "w": S.CubicBSpline.R1T1})
    g, out, ref, axes = _run(gp)
    assert out.shape == (gp.i_regular_out * gp.j_regular_out, 3, 5)
    print(f"\nRL regular worst rel L2 {_check(g, out, ref, axes):.2e}")
    g.close()


def test_rz_regular_default_points():
    gp = S.SpringsteelGridParameters(geometry="RZ", num_cells=10, iMin=0.0, iMax=100.0, kMin=0.0, kMax=50.0,
                                     kDim=16, vars={"u": 1, "v": 2}, BCB={"default": S.Chebyshev.R0,
                                                                          "v": S.Chebyshev.R1T0})
    g, out, ref, axes = _run(gp)
    assert out.shape == (gp.i_regular_out * gp.k_regular_out, 2, 5)
    print(f"\nRZ regular worst rel L2 {_check(g, out, ref, axes):.2e}")
    g.close()


def test_rlz_regular_default_points():
    gp = S.SpringsteelGridParameters(geometry="RLZ", num_cells=6, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                     kDim=8, vars={"u": 1, "v": 2})
    g, out, ref, axes = _run(gp)
    assert out.shape == (gp.i_regular_out * gp.j_regular_out * gp.k_regular_out, 2, 7)
    print(f"\nRLZ regular worst rel L2 {_check(g, out, ref, axes):.2e}")
    g.close()


def test_rlz_regular_arbitrary_points_and_device_form():
    """Unsorted, repeated and end-point coordinates; the device form gives the same bytes."""
    import torch
    gp = S.SpringsteelGridParameters(geometry="RLZ", num_cells=10, iMin=5.0, iMax=40.0, kMin=-3.0, kMax=7.0,
                                     kDim=12, vars={"u": 1, "v": 2, "w": 3})
    rng = np.random.default_rng(5)
    r = np.concatenate([[40.0, 5.0], rng.uniform(5.0, 40.0, 13)])
    lam = np.concatenate([[0.0], rng.uniform(-1.0, 7.0, 20)])
    z = np.concatenate([[-3.0, 7.0], rng.uniform(-3.0, 7.0, 9)])
    g, out, ref, axes = _run(gp, seed=2, points=(r, lam, z))
    print(f"\nRLZ arbitrary points worst rel L2 {_check(g, out, ref, axes):.2e}")
    spec = torch.from_numpy(np.ascontiguousarray(g.spectral.T)).cuda()
    dev = S.regularGridTransform_device(g, spec, r, lam, z)
    torch.cuda.synchronize()
    host = dev.cpu().numpy()                                  # (nslots, nvars, npr)
    assert np.array_equal(host, np.transpose(out, (2, 1, 0)))
    g.close()


def test_rl_regular_larger_grid():
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=60, iMin=0.0, iMax=300.0, vars={"u": 1})
    g, out, ref, axes = _run(gp, seed=4)
    print(f"\nRL nc=60 regular worst rel L2 {_check(g, out, ref, axes):.2e}")
    g.close()


def test_points_matrix_form_and_domain_error():
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=8, iMin=5.0, iMax=50.0, vars={"u": 1})
    g = S.createGrid(gp)
    pts = S.getGridpoints(g)
    g.physical[:, 0, 0] = pts[:, 0] * np.cos(pts[:, 1])
    S.spectralTransform_(g)
    reg = S.getRegularGridpoints(g)
    a = S.regularGridTransform(g, reg)
    r = np.linspace(5.0, 50.0, gp.i_regular_out)
    lam = np.array([2 * math.pi * (j - 1) / gp.j_regular_out for j in range(1, gp.j_regular_out + 1)])
    b = S.regularGridTransform(g, r, lam)
    assert np.array_equal(a, b)
    # f = r cos(lambda) is carried exactly (linear in r, wavenumber 1)
    assert np.allclose(a[:, 0, 0], reg[:, 0] * np.cos(reg[:, 1]), rtol=0, atol=1e-11)
    with pytest.raises(DomainError):
        S.regularGridTransform(g, np.array([4.0, 10.0]), lam)
    g.close()


def test_library_first_then_torch_in_fresh_process():
    """The library loaded (and its HIP runtime started) before torch touches the GPU:
    torch's own runtime must still see the device (the loader counts devices through
    torch first).  Runs in a fresh interpreter, where nothing has initialised torch."""
    import os
    import subprocess
    import sys
    code = (
        "import numpy as np, springsteel_hip as S\n"
        "gp = S.SpringsteelGridParameters(geometry='RL', num_cells=8, iMin=5.0, iMax=50.0, vars={'u': 1})\n"
        "g = S.createGrid(gp)\n"
        "pts = S.getGridpoints(g)\n"
        "g.physical[:, 0, 0] = pts[:, 0]\n"
        "S.spectralTransform_(g)\n"
        "import torch\n"
        "spec = torch.from_numpy(np.ascontiguousarray(g.spectral.T)).cuda()\n"
        "dev = S.regularGridTransform_device(g, spec, np.array([10.0, 20.0]), np.array([0.0, 1.0]))\n"
        "torch.cuda.synchronize()\n"
        "assert np.allclose(dev.cpu().numpy()[0, 0], [10.0, 10.0, 20.0, 20.0], rtol=0, atol=1e-10)\n"
        "g.close()\n"
        "print('ok')\n")
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([os.path.join(os.path.dirname(here), "springsteel.jl_amd"),
                                                         os.environ.get("PYTHONPATH", "")]))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
