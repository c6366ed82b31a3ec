"""Shared helpers of the parity tests: run one grid through the HIP library and
through the CPU oracle on the same seeded input and report relative L2 errors
per (variable, slot)."""
import os

import numpy as np

import oracle as O
import springsteel_hip as S
from springsteel_hip.metrics import rel_l2_slots

# FP64 plain form: relative L2 per (variable, slot) against the oracle
TOL_FP64 = 1e-12
# A slot whose exact value is zero (d2/dz2 of a field linear in z) holds rounding
# noise only; two FP64 implementations then agree to the noise floor of the
# derivative operator, measured against the slot's scale floor (metrics.py).
# test_oracle_pins.test_noise_floor_between_cpu_restatements measures the same
# floor between the oracle and the independent numpy restatement.
TOL_NOISE_FLOOR = 1e-9


def zero_truth_slots(ref, geometry, params, ratio=1e-6):
    """(variable, slot) pairs whose reference norm is below ratio * scale floor, i.e. slots
    that are numerically zero."""
    from springsteel_hip.metrics import slot_floors
    out = set()
    for v in range(ref.shape[1]):
        fl = slot_floors(ref[:, v, :], geometry, params)
        for s in range(ref.shape[2]):
            if fl[s] > 0 and np.linalg.norm(ref[:, v, s]) < ratio * fl[s]:
                out.add((v, s))
    return out


def rel_l2(x, ref):
    n = np.linalg.norm(ref)
    if n == 0.0:
        return float(np.linalg.norm(x))
    return float(np.linalg.norm(x - ref) / n)


def seeded_field(pts, nvars, seed, geometry):
    """Smooth seeded field on the grid points (any geometry)."""
    rng = np.random.default_rng(seed)
    out = np.empty((pts.shape[0], nvars), order="F")
    x = pts[:, 0]
    span = max(np.ptp(x), 1e-30)
    xs = (x - x.min()) / span
    for v in range(nvars):
        c = rng.normal(size=8)
        f = c[0] + c[1] * xs + c[2] * np.sin(3 * xs) + c[3] * np.exp(-4 * xs * xs)
        if geometry in ("RL", "RLZ", "SL", "SLZ"):
            lam = pts[:, 1]
            f = f + xs * (c[4] * np.cos(lam) + c[5] * np.sin(2 * lam) + 0.3 * c[6] * np.cos(5 * lam))
        if geometry in ("RZ", "RLZ", "SLZ"):
            z = pts[:, -1]
            zs = (z - z.min()) / max(np.ptp(z), 1e-30)
            f = f * (1.0 + 0.5 * c[7] * zs) + 0.2 * np.cos(3 * zs) * xs
        out[:, v] = f
    return out


def sensitivity(og, physical, ophys, geometry, params, seed=123):
    """Per (variable, slot) change of the oracle's output when its input is perturbed at
    the rounding level (relative 2**-52, random signs), measured with the slot floors.
    Second derivatives amplify rounding by the square of the derivative operator's norm,
    so two correct FP64 implementations can only agree to about this level."""
    rng = np.random.default_rng(seed)
    pert = np.asfortranarray(physical.copy())
    x = pert[:, :, 0]
    pert[:, :, 0] = x * (1.0 + 2.0 ** -52 * rng.choice([-1.0, 1.0], size=x.shape))
    th = min(len(os.sched_getaffinity(0)), 16)
    p2 = og.grid_transform(og.spectral_transform(pert, threads=th), threads=th)
    return [rel_l2_slots(p2[:, v, :], ophys[:, v, :], geometry, params) for v in range(ophys.shape[1])]


def slot_tolerances(sens, tol=None):
    """max(TOL_FP64, SENS_FACTOR * sensitivity) per (variable, slot)."""
    tol = TOL_FP64 if tol is None else tol
    return [[max(tol, SENS_FACTOR * s) for s in row] for row in sens]


# margin over the measured rounding sensitivity of the oracle: the worst GPU error
# measured is 8.7 times it (d2/dz2 of fields linear in z, composed inverse Chebyshev
# matrices against the reference's recurrences)
SENS_FACTOR = 12.0


def run_both(gp, field=None, seed=0):
    """Returns (grid, oracle_grid, gpu_spectral, ora_spectral, gpu_physical, ora_physical,
    gpu_physical_from_ora_spectral, input_physical)."""
    grid = S.createGrid(gp)
    og = O.OracleGrid(grid.resolved)
    pts = S.getGridpoints(grid)
    opts = og.gridpoints()
    assert pts.shape == opts.shape and np.allclose(pts, opts, rtol=0, atol=1e-9 * max(1.0, np.abs(opts).max()))
    nv = grid.resolved.nvars
    f = seeded_field(pts, nv, seed, grid.geometry) if field is None else field(pts, nv)
    grid.physical[:, :, 0] = f
    inp = np.asfortranarray(grid.physical.copy())
    S.spectralTransform_(grid)
    ospec = og.spectral_transform(inp, threads=8)
    S.gridTransform_(grid)
    ophys = og.grid_transform(ospec, threads=8)
    # inverse alone, on the oracle's spectral input
    phys2 = np.zeros_like(grid.physical, order="F")
    S.gridTransform(grid, phys2, np.asfortranarray(ospec))
    return grid, og, grid.spectral.copy(), ospec, grid.physical.copy(), ophys, phys2, inp


def errors(a, b, geometry=None, params=None):
    """rel L2 per variable (spectral) or per (variable, slot) (physical, with the slot scale
    floors of springsteel_hip.metrics)."""
    if a.ndim == 2:
        return [rel_l2(a[:, v], b[:, v]) for v in range(a.shape[1])]
    if geometry is None:
        return [[rel_l2(a[:, v, s], b[:, v, s]) for s in range(a.shape[2])] for v in range(a.shape[1])]
    return [rel_l2_slots(a[:, v, :], b[:, v, :], geometry, params) for v in range(a.shape[1])]
