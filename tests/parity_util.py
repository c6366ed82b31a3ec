"""Shared helpers of the parity tests: run one grid through the HIP library and
through the CPU oracle on the same seeded input and report relative L2 errors
per (variable, slot)."""
import numpy as np

import oracle as O
import springsteel_hip as S

# FP64 plain form: relative L2 per (variable, slot) against the oracle
TOL_FP64 = 1e-12


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
        if geometry in ("RL", "RLZ"):
            lam = pts[:, 1]
            f = f + xs * (c[4] * np.cos(lam) + c[5] * np.sin(2 * lam) + 0.3 * c[6] * np.cos(5 * lam))
        if geometry in ("RZ", "RLZ"):
            z = pts[:, -1]
            zs = (z - z.min()) / max(np.ptp(z), 1e-30)
            f = f * (1.0 + 0.5 * c[7] * zs) + 0.2 * np.cos(3 * zs) * xs
        out[:, v] = f
    return out


def run_both(gp, field=None, seed=0):
    """Returns (grid, oracle_grid, gpu_spectral, ora_spectral, gpu_physical, ora_physical,
    gpu_physical_from_ora_spectral)."""
    grid = S.createGrid(gp)
    og = O.OracleGrid(grid.resolved)
    pts = S.getGridpoints(grid)
    opts = og.gridpoints()
    assert pts.shape == opts.shape and np.allclose(pts, opts, rtol=0, atol=1e-9 * max(1.0, np.abs(opts).max()))
    nv = grid.resolved.nvars
    f = seeded_field(pts, nv, seed, grid.geometry) if field is None else field(pts, nv)
    grid.physical[:, :, 0] = f
    S.spectralTransform_(grid)
    ospec = og.spectral_transform(np.asfortranarray(grid.physical), threads=8)
    S.gridTransform_(grid)
    ophys = og.grid_transform(ospec, threads=8)
    # inverse alone, on the oracle's spectral input
    phys2 = np.zeros_like(grid.physical, order="F")
    S.gridTransform(grid, phys2, np.asfortranarray(ospec))
    return grid, og, grid.spectral.copy(), ospec, grid.physical.copy(), ophys, phys2


def errors(a, b):
    """max over variables of rel L2 per slot (physical) or per variable (spectral)."""
    if a.ndim == 2:
        return [rel_l2(a[:, v], b[:, v]) for v in range(a.shape[1])]
    return [[rel_l2(a[:, v, s], b[:, v, s]) for s in range(a.shape[2])] for v in range(a.shape[1])]
