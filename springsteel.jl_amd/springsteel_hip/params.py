"""Grid parameters of the reference API (SpringsteelGridParameters,
src/Springsteel.jl:425-465) and their per-variable resolution
(src/factory.jl:166-175, 290-336, 646-740)."""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Dict, List

from .bcs import (CubicBSpline, Chebyshev, Fourier, chebyshev_bc_kind, spline_bc_kind)
from .errors import ArgumentError, DomainError, NotSupportedError

# geometry aliases of src/factory.jl:127-140 restricted to RL / RZ / RLZ / SL / SLZ
_GEOMETRY_ALIASES = {"RL": "RL", "Polar": "RL", "RZ": "RZ", "RLZ": "RLZ", "Cylindrical": "RLZ",
                     "SL": "SL", "SphericalShell": "SL", "SLZ": "SLZ", "Sphere": "SLZ"}
GEOMETRY_CODES = {"RL": 1, "RZ": 2, "RLZ": 3, "SL": 4, "SLZ": 5}

_GL = {1: ([0.0], [2.0]),
       2: ([-1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0)], [1.0, 1.0]),
       3: ([-math.sqrt(3.0 / 5.0), 0.0, math.sqrt(3.0 / 5.0)], [5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])}


def quadpoints(mubar: int, quadrature: str):
    """Cell-local quadrature positions (src/CubicBSpline.jl:172-200)."""
    if quadrature == "regular":
        return [(2 * k - 1) / (2 * mubar) for k in range(1, mubar + 1)]
    if mubar in _GL:
        nodes = _GL[mubar][0]
    elif mubar == 4:
        a = math.sqrt(3.0 / 7.0 + 2.0 / 7.0 * math.sqrt(6.0 / 5.0))
        b = math.sqrt(3.0 / 7.0 - 2.0 / 7.0 * math.sqrt(6.0 / 5.0))
        nodes = [-a, -b, b, a]
    elif mubar == 5:
        a = math.sqrt(5.0 + 2.0 * math.sqrt(10.0 / 7.0)) / 3.0
        b = math.sqrt(5.0 - 2.0 * math.sqrt(10.0 / 7.0)) / 3.0
        nodes = [-a, -b, 0.0, b, a]
    else:
        raise ArgumentError(f"Gauss-Legendre tables only available for mubar 1-5, got {mubar}")
    return [(x + 1.0) / 2.0 for x in nodes]


def mish_points(gp) -> List[float]:
    """Spline mish points x = iMin + mc DX + q_mu DX (src/CubicBSpline.jl:887-897)."""
    DX = (gp.iMax - gp.iMin) / gp.num_cells
    q = quadpoints(gp.mubar, gp.quadrature)
    return [gp.iMin + mc * DX + q[mu] * DX for mc in range(gp.num_cells) for mu in range(gp.mubar)]


def ring_layout(gp) -> List[tuple]:
    """(points N, wavenumbers of the ring kr, phase offset ymin) of every ring, 1-based ring r.
    Cylindrical (src/factory.jl:153-156, 608-624): ri = r + patchOffsetL, N = 4 + 4 ri,
    kr = ri, ymin = dl (ri - 1) / 2.  Spherical (159-163, 626-643): kr = max(1,
    round(sin(theta_r) max_ri)) (ties to even, as Julia's round), N = 4 kr, ymin = dl r / 2."""
    geom = normalize_geometry(gp.geometry)
    out = []
    if geom in ("SL", "SLZ"):
        max_ri = gp.iDim + gp.patchOffsetL
        for r, th in enumerate(mish_points(gp), start=1):
            kr = max(1, int(round(math.sin(th) * max_ri)))
            n = 4 * kr
            out.append((n, kr, 0.5 * (2 * math.pi / n) * r))
    else:
        for r in range(1, gp.iDim + 1):
            ri = r + gp.patchOffsetL
            n = 4 + 4 * ri
            out.append((n, ri, 0.5 * (2 * math.pi / n) * (ri - 1)))
    return out


@dataclass
class SpringsteelGridParameters:
    """Mirror of the reference's ``Base.@kwdef struct SpringsteelGridParameters``.

    Derived defaults (iDim, b_iDim, b_kDim, patchOffsetL, ...) are filled in by
    ``__post_init__`` when left at ``None``, with the reference's formulas.
    """
    geometry: str = "1D"
    iMin: float = 0.0
    iMax: float = 0.0
    num_cells: int = 0
    mubar: int = 3
    quadrature: str = "gauss"
    iDim: int = None
    b_iDim: int = None
    l_q: Dict = field(default_factory=lambda: {"default": 2.0})
    BCL: Dict = field(default_factory=lambda: {"default": CubicBSpline.R0})
    BCR: Dict = field(default_factory=lambda: {"default": CubicBSpline.R0})
    jMin: float = 0.0
    jMax: float = 2 * math.pi
    max_wavenumber: Dict = field(default_factory=lambda: {"default": -1})
    jDim: int = 0
    b_jDim: int = 0
    BCU: Dict = field(default_factory=lambda: {"default": Fourier.PERIODIC})
    BCD: Dict = field(default_factory=lambda: {"default": Fourier.PERIODIC})
    kMin: float = 0.0
    kMax: float = 0.0
    kDim: int = 0
    b_kDim: int = None
    BCB: Dict = field(default_factory=lambda: {"default": Chebyshev.R0})
    BCT: Dict = field(default_factory=lambda: {"default": Chebyshev.R0})
    vars: Dict = field(default_factory=lambda: {"u": 1})
    fourier_filter: Dict = field(default_factory=dict)
    chebyshev_filter: Dict = field(default_factory=dict)
    spectralIndexL: int = 1
    spectralIndexR: int = None
    patchOffsetL: int = None
    patchOffsetR: int = None
    tile_num: int = 0
    i_regular_out: int = None
    j_regular_out: int = None
    k_regular_out: int = None

    def __post_init__(self):
        if self.iDim is None:
            self.iDim = self.num_cells * self.mubar
        if self.b_iDim is None:
            self.b_iDim = self.num_cells + 3
        if self.b_kDim is None:
            self.b_kDim = min(self.kDim, (2 * self.kDim - 1) // 3 + 1)
        if self.spectralIndexR is None:
            self.spectralIndexR = self.spectralIndexL + self.b_iDim - 1
        if self.patchOffsetL is None:
            self.patchOffsetL = (self.spectralIndexL - 1) * self.mubar
        if self.patchOffsetR is None:
            self.patchOffsetR = self.patchOffsetL + self.iDim
        # regular-grid output sizes (src/Springsteel.jl:461-464)
        if self.i_regular_out is None:
            self.i_regular_out = self.num_cells + 1
        if self.j_regular_out is None:
            self.j_regular_out = self.iDim * 2 + 1
        if self.k_regular_out is None:
            self.k_regular_out = self.kDim + 1


def normalize_geometry(g: str) -> str:
    return _GEOMETRY_ALIASES.get(g, g)


def _get(d: Dict, key, fallback=None):
    """Per-variable lookup with "default" fallback (src/factory.jl:329-336)."""
    if key in d:
        return d[key]
    if "default" in d:
        return d["default"]
    if fallback is not None:
        return fallback
    raise ArgumentError(f"no entry for variable {key!r} and no 'default'")


def compute_derived_params(gp: SpringsteelGridParameters) -> SpringsteelGridParameters:
    """src/factory.jl:290-324: ring dims sum(N), sum(1 + 2 kr) for RL / RLZ / SL / SLZ
    (_fourier_j_dims_cyl 166-175, _fourier_j_dims_sph 193-203)."""
    geom = normalize_geometry(gp.geometry)
    if geom in ("RL", "RLZ", "SL", "SLZ"):
        rings = ring_layout(gp)
        return replace(gp, jDim=sum(n for n, _, _ in rings), b_jDim=sum(1 + 2 * kr for _, kr, _ in rings))
    return gp


@dataclass
class ResolvedGrid:
    """Everything the C-ABI descriptor needs, resolved per variable (0-based var index)."""
    geometry: int
    num_cells: int
    mubar: int
    quadrature: int
    iMin: float
    iMax: float
    nvars: int
    l_q: List[float]
    bcl: List[tuple]
    bcr: List[tuple]
    kMin: float
    kMax: float
    kDim: int
    b_kDim: int
    bcb: List[int]
    bct: List[int]
    max_wavenumber: List[int]
    ring_kmax: int
    patchOffsetL: int
    has_filters: bool


def resolve(gp: SpringsteelGridParameters) -> ResolvedGrid:
    geom = normalize_geometry(gp.geometry)
    if geom not in GEOMETRY_CODES:
        raise DomainError(f"geometry {gp.geometry!r} is not an RL, RZ, RLZ, SL or SLZ grid")
    code = GEOMETRY_CODES[geom]
    if gp.quadrature not in ("gauss", "regular"):
        raise ArgumentError(f"Unknown quadrature type: {gp.quadrature}")
    names = sorted(gp.vars, key=lambda k: gp.vars[k])
    idx = [gp.vars[k] for k in names]
    if idx != list(range(1, len(idx) + 1)):
        raise ArgumentError("vars must map names to 1..nvars")
    l_q, bcl, bcr, bcb, bct, mw = [], [], [], [], [], []
    for key in names:
        l_q.append(float(_get(gp.l_q, key, 2.0)))
        bcl.append(spline_bc_kind(_get(gp.BCL, key)))
        bcr.append(spline_bc_kind(_get(gp.BCR, key)))
        if code in (2, 3, 5):
            bcb.append(chebyshev_bc_kind(_get(gp.BCB, key)))
            bct.append(chebyshev_bc_kind(_get(gp.BCT, key)))
        else:
            bcb.append(0)
            bct.append(0)
        mw.append(int(_get(gp.max_wavenumber, key, -1)))
    ring_kmax = -1
    if code in (3, 5):
        # RLZ / SLZ rings are shared by all variables; the reference takes the cap
        # of the last key of its Dict iteration (src/factory.jl:731-734, 828-831),
        # so the cap is only well defined when every variable resolves to the same value.
        caps = set(mw)
        if len(caps) > 1:
            raise NotSupportedError(f"{geom} variables with different max_wavenumber: the reference's choice "
                                    "depends on Julia Dict iteration order")
        ring_kmax = mw[0]
    has_filters = bool(gp.fourier_filter) or bool(gp.chebyshev_filter)
    return ResolvedGrid(code, gp.num_cells, gp.mubar, 0 if gp.quadrature == "gauss" else 1, float(gp.iMin),
                        float(gp.iMax), len(names), l_q, bcl, bcr, float(gp.kMin), float(gp.kMax), int(gp.kDim),
                        int(gp.b_kDim), bcb, bct, mw, ring_kmax, int(gp.patchOffsetL), has_filters)
