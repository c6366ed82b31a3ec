"""Named synthetic workloads (grid parameters + seeded input fields).

  rl_vortex     RL, radius 0-500, 100 cells (300 radial points, rings 4+4r),
                natural BCs, 1 variable: Gaussian vortex exp(-(r/50)^2) plus
                Fourier modes k <= 8 with N(0, 0.1^2) amplitudes.
  rz_modes      RZ, 256 cells on [0, 1e5] x 64 Chebyshev points on [0, 20000],
                4 variables: sum_{k<=10} a_k sin(k pi x / L) T_k(xi), a_k ~ N(0, 1/k^2).
  rlz_vortex    RLZ, 200 cells on [0, 500], rings 4+4r up to 2404 points,
                32 Chebyshev points on [0, 20000], 5 variables: Rankine-like
                vortex plus azimuthal modes k <= 16 with N(0, 0.05^2)
                amplitudes, varying linearly in z.
  slz_sphere    SLZ, 200 cells on colatitude [0, pi], rings of 4 round(600 sin theta)
                points, 32 Chebyshev points on [0, 20000], 5 variables: sin(theta)
                times azimuthal modes k <= 16 with N(0, 0.05^2) amplitudes, varying
                linearly in z.
Seeds come from numpy.random.default_rng so every run sees the same fields.
"""
from __future__ import annotations

import numpy as np

from .params import SpringsteelGridParameters


def rl_params(num_cells=100, iMax=500.0, nvars=1):
    return SpringsteelGridParameters(geometry="RL", iMin=0.0, iMax=iMax, num_cells=num_cells,
                                     vars={f"v{i}": i + 1 for i in range(nvars)})


def rz_params(num_cells=256, iMax=1.0e5, kDim=64, kMax=20000.0, nvars=4):
    return SpringsteelGridParameters(geometry="RZ", iMin=0.0, iMax=iMax, num_cells=num_cells, kMin=0.0,
                                     kMax=kMax, kDim=kDim, vars={f"v{i}": i + 1 for i in range(nvars)})


def rlz_params(num_cells=200, iMax=500.0, kDim=32, kMax=20000.0, nvars=5):
    return SpringsteelGridParameters(geometry="RLZ", iMin=0.0, iMax=iMax, num_cells=num_cells, kMin=0.0,
                                     kMax=kMax, kDim=kDim, vars={f"v{i}": i + 1 for i in range(nvars)})


def slz_params(num_cells=200, kDim=32, kMax=20000.0, nvars=5):
    return SpringsteelGridParameters(geometry="SLZ", iMin=0.0, iMax=float(np.pi), num_cells=num_cells, kMin=0.0,
                                     kMax=kMax, kDim=kDim, vars={f"v{i}": i + 1 for i in range(nvars)})


WORKLOADS = {
    "rl_vortex": dict(params=rl_params, desc="RL nc=100 r in [0,500], 1 var"),
    "rz_modes": dict(params=rz_params, desc="RZ nc=256 x kDim=64, 4 vars"),
    "rlz_vortex": dict(params=rlz_params, desc="RLZ nc=200 x kDim=32, 5 vars"),
    "slz_sphere": dict(params=slz_params, desc="SLZ nc=200 x kDim=32, 5 vars"),
}


def rl_field(pts: np.ndarray, nvars: int, seed: int = 0) -> np.ndarray:
    """pts (npts, 2) -> field (npts, nvars)."""
    r, lam = pts[:, 0], pts[:, 1]
    out = np.empty((len(r), nvars), order="F")
    for v in range(nvars):
        rng = np.random.default_rng(seed + v)
        A = rng.normal(0.0, 0.1, 8)
        B = rng.normal(0.0, 0.1, 8)
        f = np.exp(-(r / 50.0) ** 2)
        for k in range(1, 9):
            f = f + A[k - 1] * np.cos(k * lam) + B[k - 1] * np.sin(k * lam)
        out[:, v] = f
    return out


def rz_field(pts: np.ndarray, nvars: int, L: float, zmax: float, seed: int = 1) -> np.ndarray:
    x, z = pts[:, 0], pts[:, 1]
    xi = 2.0 * z / zmax - 1.0
    out = np.empty((len(x), nvars), order="F")
    for v in range(nvars):
        rng = np.random.default_rng(seed + v)
        a = rng.normal(0.0, 1.0, 10) / np.arange(1, 11)
        f = np.zeros_like(x)
        for k in range(1, 11):
            f += a[k - 1] * np.sin(k * np.pi * x / L) * np.cos(k * np.arccos(np.clip(xi, -1.0, 1.0)))
        out[:, v] = f
    return out


def rlz_field_var(r: np.ndarray, lam: np.ndarray, z: np.ndarray, v: int, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed + v)
    A = rng.normal(0.0, 0.05, 16)
    B = rng.normal(0.0, 0.05, 16)
    s = r / 50.0
    f = 2.0 * s / (1.0 + s * s)
    for k in range(1, 17):
        f = f + A[k - 1] * np.cos(k * lam) + B[k - 1] * np.sin(k * lam)
    return f * (1.0 - z / 40000.0)


def slz_field_var(theta: np.ndarray, lam: np.ndarray, v: int, seed: int = 3) -> np.ndarray:
    """The (theta, lambda) part of the SLZ workload's variable v (the z factor is applied by the caller)."""
    rng = np.random.default_rng(seed + v)
    A = rng.normal(0.0, 0.05, 16)
    B = rng.normal(0.0, 0.05, 16)
    f = np.ones_like(theta)
    for k in range(1, 17):
        f = f + A[k - 1] * np.cos(k * lam) + B[k - 1] * np.sin(k * lam)
    return np.sin(theta) * f


def rlz_field(pts: np.ndarray, nvars: int, seed: int = 2) -> np.ndarray:
    out = np.empty((pts.shape[0], nvars), order="F")
    for v in range(nvars):
        out[:, v] = rlz_field_var(pts[:, 0], pts[:, 1], pts[:, 2], v, seed)
    return out


def workload_input(name: str, grid) -> np.ndarray:
    """Seeded input of a named workload on a grid: (npts, nvars) in the physical layout.
    The ring grids are built separably: (r or theta, lambda) of every ring point from
    the 2-D grid of the same rings, times the z factor of one column."""
    from .grid import SpringsteelGrid, getGridpoints
    from .params import SpringsteelGridParameters, compute_derived_params
    p = grid.params
    nv = grid.resolved.nvars
    if name in ("slz_sphere", "rlz_vortex"):
        geo2 = "SL" if name == "slz_sphere" else "RL"
        g2 = SpringsteelGrid(compute_derived_params(
            SpringsteelGridParameters(**{**p.__dict__, "geometry": geo2, "jDim": 0, "b_jDim": 0})))
        pts2 = getGridpoints(g2)
        scale, offset = -0.5 * (p.kMax - p.kMin), 0.5 * (p.kMin + p.kMax)
        zs = np.cos(np.arange(p.kDim) * np.pi / (p.kDim - 1)) * scale + offset
        out = np.empty((grid.npts, nv), order="F")
        for v in range(nv):
            if name == "slz_sphere":
                base = slz_field_var(pts2[:, 0], pts2[:, 1], v)
            else:
                base = rlz_field_var(pts2[:, 0], pts2[:, 1], np.zeros(len(pts2)), v)
            out[:, v] = np.outer(base, 1.0 - zs / 40000.0).ravel()
        return out
    pts = getGridpoints(grid)
    if name == "rz_modes":
        return rz_field(pts, nv, L=p.iMax, zmax=p.kMax)
    return rl_field(pts, nv)


# --- batching workload ---------------------------------------------------------------------------
# 64 independent grids: num_cells ~ U{50..400}, Chebyshev points ~ {1, 16, 32, 64} (1 -> RL grid),
# 1..6 variables, fields as rl_field (constant in z) with per-grid seeds.

def grid_device_bytes(nc: int, nz: int, nv: int, fp32: bool = False) -> int:
    """Device bytes of one ring grid: physical + spectral arrays and the plan's intermediates."""
    iDim = 3 * nc
    K = iDim
    jDim = 2 * iDim * iDim + 6 * iDim            # sum_{r=1..iDim} (4 + 4r)
    b_jDim = iDim * iDim + 2 * iDim              # sum (2r + 1)
    Ktot = iDim * (iDim + 1) // 2 + iDim         # sum (r + 1)
    b_iDim = nc + 3
    if nz == 1:
        nzb, npts, nslots = 1, jDim, 5
    else:
        nzb, npts, nslots = min(nz, (2 * nz - 1) // 3 + 1), jDim * nz, 7
    nspec = nzb * b_iDim * (1 + 2 * K)
    ntr = nv * nzb
    e = 4 if fp32 else 8
    # rc, X, cm (RLZ), the solve scratch and the chirp-z synthesis output Y (RLZ)
    inter = ntr * b_jDim * e + 6 * ntr * Ktot * e + (ntr * jDim * e if nz > 1 else 0) + b_iDim * 2 * ntr * (K + 1) * 8
    if nz > 1:
        inter += 5 * ntr * jDim * e
    return nslots * nv * npts * 8 + nv * nspec * 8 + inter


def batch_params(ngrids: int = 64, seed: int = 0, budget_bytes: float = 200e9):
    """Grids of the batching workload: ngrids draws from the distribution above.  The
    literal draw needs 360-600 GB, more than one GPU holds, so when its FP64 footprint
    exceeds budget_bytes every num_cells is scaled by one common factor
    sqrt(budget / footprint) (footprints grow as num_cells**2), keeping the spread of
    sizes, the Chebyshev dimensions and the variable counts of the draw.
    Returns (list of SpringsteelGridParameters, field seeds, FP64 bytes, scale)."""
    rng = np.random.default_rng(seed)
    nc = rng.integers(50, 401, ngrids)
    nz = rng.choice([1, 16, 32, 64], ngrids)
    nv = rng.integers(1, 7, ngrids)
    total = sum(grid_device_bytes(int(a), int(b), int(c)) for a, b, c in zip(nc, nz, nv))
    scale = min(1.0, float(np.sqrt(budget_bytes / total)))
    while True:
        ncs = [max(8, int(round(a * scale))) for a in nc]
        total = sum(grid_device_bytes(a, int(b), int(c)) for a, b, c in zip(ncs, nz, nv))
        if total <= budget_bytes:
            break
        scale *= 0.98
    out = []
    for g, (a, b, c) in enumerate(zip(ncs, nz, nv)):
        vars_ = {f"v{i}": i + 1 for i in range(int(c))}
        if b == 1:
            out.append(SpringsteelGridParameters(geometry="RL", iMin=0.0, iMax=500.0, num_cells=a, vars=vars_))
        else:
            out.append(SpringsteelGridParameters(geometry="RLZ", iMin=0.0, iMax=500.0, num_cells=a, kMin=0.0,
                                                 kMax=20000.0, kDim=int(b), vars=vars_))
    return out, [1000 + g for g in range(ngrids)], total, scale
