"""Boundary conditions of the grid API, mirrored from the reference.

``BoundaryConditions`` and its constructors follow src/Springsteel.jl:45-151;
the legacy Dict forms follow src/CubicBSpline.jl:47-118 (spline) and
src/Chebyshev.jl:35-96 (Chebyshev).  Conversion to the C-ABI kinds follows
src/factory.jl:343-396 (``_bc_to_spline_dict`` / ``_bc_to_chebyshev_dict``)
and the GammaBC dispatch of src/CubicBSpline.jl:380-420.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

from .errors import ArgumentError, DomainError

# C-ABI spline BC kinds (include/springsteel_hip.h)
SSG_BC_R0, SSG_BC_ALPHA1, SSG_BC_ALPHA2, SSG_BC_PERIODIC, SSG_BC_R3, SSG_BC_R3X = range(6)
SSG_CBC_R0, SSG_CBC_R1T0, SSG_CBC_R1T1, SSG_CBC_OTHER = range(4)


@dataclass(frozen=True)
class BoundaryConditions:
    """src/Springsteel.jl:45-63."""
    u: Optional[float] = None
    du: Optional[float] = None
    d2u: Optional[float] = None
    robin: Optional[Tuple[float, float, float]] = None
    periodic: bool = False

    def __post_init__(self):
        if self.robin is not None and any(x is not None for x in (self.u, self.du, self.d2u)):
            raise ArgumentError("Robin BC is mutually exclusive with u/du/d2u constraint slots")
        if self.periodic and any(x is not None for x in (self.u, self.du, self.d2u, self.robin)):
            raise ArgumentError("PeriodicBC must have all other slots as nothing")


def NaturalBC():
    return BoundaryConditions()


def DirichletBC(v=0.0):
    return BoundaryConditions(u=float(v))


def NeumannBC(v=0.0):
    return BoundaryConditions(du=float(v))


def SecondDerivativeBC(v=0.0):
    return BoundaryConditions(d2u=float(v))


def RobinBC(alpha, beta, gamma=0.0):
    return BoundaryConditions(robin=(float(alpha), float(beta), float(gamma)))


def PeriodicBC():
    return BoundaryConditions(periodic=True)


def CauchyBC(u, du):
    return BoundaryConditions(u=float(u), du=float(du))


def ExponentialBC(lam):
    return RobinBC(1.0, -float(lam), 0.0)


def SymmetricBC():
    return NeumannBC(0.0)


def AntisymmetricBC():
    return BoundaryConditions(u=0.0, d2u=0.0)


def ZerosBC():
    return BoundaryConditions(u=0.0, du=0.0, d2u=0.0)


def FixedBC(u=math.nan, du=0.0, d2u=0.0):
    if math.isnan(u):
        return BoundaryConditions(u=math.nan, du=math.nan, d2u=math.nan)
    return BoundaryConditions(u=float(u), du=float(du), d2u=float(d2u))


def bc_rank(bc: BoundaryConditions) -> int:
    """src/Springsteel.jl:147-151."""
    if bc.periodic:
        return 0
    if bc.robin is not None:
        return 1
    return sum(x is not None for x in (bc.u, bc.du, bc.d2u))


def is_inhomogeneous(bc: BoundaryConditions) -> bool:
    if bc.robin is not None:
        return bc.robin[2] != 0.0
    return any(x is not None and x != 0.0 for x in (bc.u, bc.du, bc.d2u))


class CubicBSpline:
    """Legacy spline BC dicts (src/CubicBSpline.jl:47-118)."""
    R0 = {"R0": 0}
    R1T0 = {"α1": -4.0, "β1": -1.0}
    R1T1 = {"α1": 0.0, "β1": 1.0}
    R1T2 = {"α1": 2.0, "β1": -1.0}
    R2T10 = {"α2": 1.0, "β2": -0.5}
    R2T20 = {"α2": -1.0, "β2": 0.0}
    R3 = {"R3": 0}
    R3X = {"R3X": 0}
    PERIODIC = {"PERIODIC": 0}


class Chebyshev:
    """Legacy Chebyshev BC dicts (src/Chebyshev.jl:35-96)."""
    R0 = {"R0": 0}
    R1T0 = {"α0": 0.0}
    R1T1 = {"α1": 0.0}
    R1T2 = {"α2": 0.0}
    R2T10 = {"β1": 0.0, "β2": 0.0}
    R2T20 = {"β1": 0.0, "β2": 0.0}
    R3 = {"R3": 0}


class Fourier:
    PERIODIC = {"PERIODIC": 0}


def spline_bc_dict(bc) -> dict:
    """BoundaryConditions -> internal spline Dict (src/factory.jl:343-375)."""
    if isinstance(bc, dict):
        return bc
    if bc.periodic:
        return CubicBSpline.PERIODIC
    r = bc_rank(bc)
    if r == 0:
        return CubicBSpline.R0
    if r == 1:
        if bc.u is not None:
            return {"α1": -4.0, "β1": -1.0}
        if bc.du is not None:
            return {"α1": 0.0, "β1": 1.0}
        if bc.d2u is not None:
            return {"α1": 2.0, "β1": -1.0}
        a, b, _ = bc.robin
        lam = -b / a
        return {"α1": -4.0 / (3 * lam + 1), "β1": (3 * lam - 1) / (3 * lam + 1)}
    if r == 2:
        if bc.u is not None and bc.du is not None:
            return {"α2": 1.0, "β2": -0.5}
        if bc.u is not None and bc.d2u is not None:
            return {"α2": -1.0, "β2": 0.0}
        raise ArgumentError("Unsupported rank-2 BC combination for CubicBSpline: du + d2u without u is not defined")
    return {"R3X": 0} if is_inhomogeneous(bc) else {"R3": 0}


def spline_bc_kind(bc) -> Tuple[int, float, float]:
    """Internal spline Dict -> (C-ABI kind, alpha, beta), in GammaBC's order of checks."""
    d = spline_bc_dict(bc)
    if "α1" in d:
        return SSG_BC_ALPHA1, float(d["α1"]), float(d["β1"])
    if "α2" in d:
        return SSG_BC_ALPHA2, float(d["α2"]), float(d["β2"])
    if d == CubicBSpline.R0:
        return SSG_BC_R0, 0.0, 0.0
    if d == CubicBSpline.R3:
        return SSG_BC_R3, 0.0, 0.0
    if "R3X" in d:
        return SSG_BC_R3X, 0.0, 0.0
    if d == CubicBSpline.PERIODIC:
        return SSG_BC_PERIODIC, 0.0, 0.0
    raise ArgumentError(f"GammaBC: unrecognised BC: {d}")


def chebyshev_bc_dict(bc) -> dict:
    """BoundaryConditions -> internal Chebyshev Dict (src/factory.jl:378-396)."""
    if isinstance(bc, dict):
        return bc
    if bc.periodic:
        raise ArgumentError("PeriodicBC is not supported for Chebyshev basis")
    if bc.robin is not None:
        raise ArgumentError("RobinBC is not supported for Chebyshev basis")
    r = bc_rank(bc)
    if r == 0:
        return Chebyshev.R0
    if r == 1:
        if bc.u is not None:
            return {"α0": bc.u}
        if bc.du is not None:
            return {"α1": bc.du}
        return {"α2": bc.d2u}
    raise ArgumentError(f"Rank-{r} BCs are not yet implemented for Chebyshev basis")


def chebyshev_bc_kind(bc) -> int:
    """src/Chebyshev.jl:949-963 (_bc_type): classification by key."""
    d = chebyshev_bc_dict(bc)
    if "R0" in d:
        return SSG_CBC_R0
    if "α0" in d:
        return SSG_CBC_R1T0
    if "α1" in d:
        return SSG_CBC_R1T1
    return SSG_CBC_OTHER


def require_domain(kind: int):
    if kind == SSG_CBC_OTHER:
        raise DomainError("Unrecognized Chebyshev BC dict keys")
