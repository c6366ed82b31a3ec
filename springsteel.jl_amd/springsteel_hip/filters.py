"""Spectral filters: the reference's filter types and weights (src/filtering.jl).

The weights are host arithmetic on scalars; the library applies them to the
spectral lines on the GPU (ssg_plan_set_filter / ssg_apply_filter).

  SpectralFilter   src/filtering.jl:67-73  (low_pass, high_pass, notch, window, taper_width)
  GaussianFilter   src/filtering.jl:100-103 (sigma, order)
  window_weight    _window_weight, src/filtering.jl:113-133
  filter_weight    _filter_weight, src/filtering.jl:144-187
  get_filter       _get_filter, src/filtering.jl:199-203 (variable key, then "default")
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .errors import ArgumentError

_RTOL = math.sqrt(np.finfo(np.float64).eps)   # Julia's isapprox default rtol


def isapprox(x: float, y: float) -> bool:
    """Julia's isapprox(x, y) with its defaults (atol 0, rtol sqrt(eps))."""
    return x == y or (math.isfinite(x) and math.isfinite(y) and abs(x - y) <= _RTOL * max(abs(x), abs(y)))


@dataclass
class SpectralFilter:
    """Wavenumber-domain filter with optional windowed taper (src/filtering.jl:67-73)."""
    low_pass: int = -1
    high_pass: int = 0
    notch: List[int] = field(default_factory=list)
    window: str = "boxcar"
    taper_width: int = 0


@dataclass
class GaussianFilter:
    """exp(-(k / sigma)^(2 order)) (src/filtering.jl:100-103)."""
    sigma: float
    order: int = 1


def window_weight(window: str, t: float) -> float:
    """_window_weight (src/filtering.jl:113-133): t = 0 passband edge, t = 1 stopband edge."""
    t = min(max(float(t), 0.0), 1.0)
    if window == "boxcar":
        return 1.0 if t < 1.0 else 0.0
    if window == "hann":
        return 0.5 * (1.0 + math.cos(math.pi * t))
    if window == "lanczos":
        if isapprox(t, 0.0):
            return 1.0
        return math.sin(math.pi * t) / (math.pi * t)
    if window == "exponential":
        return math.exp(-t ** 2 * 4.0)
    raise ArgumentError(f"Unknown window function: {window}. Use :boxcar, :hann, :lanczos, or :exponential.")


def _taper(f: SpectralFilter, dist: int) -> float:
    if f.taper_width > 0:
        if dist > f.taper_width:
            return 0.0
        return window_weight(f.window, float(dist) / float(f.taper_width))
    return 0.0


def filter_weight(f, k: int) -> float:
    """_filter_weight (src/filtering.jl:144-187): the weight of wavenumber / mode k."""
    if isinstance(f, GaussianFilter):
        return math.exp(-(float(k) / f.sigma) ** (2 * f.order))
    if not isinstance(f, SpectralFilter):
        raise ArgumentError(f"not a spectral filter: {f!r}")
    if k in f.notch:
        return 0.0
    if f.high_pass > 0 and k < f.high_pass:
        return _taper(f, f.high_pass - k)
    if f.low_pass >= 0 and k > f.low_pass:
        return _taper(f, k - f.low_pass)
    return 1.0


def get_filter(filters: dict, name: str):
    """_get_filter (src/filtering.jl:199-203)."""
    if name in filters:
        return filters[name]
    return filters.get("default")


def filter_weights(gp, names, K: int, b_kDim: int):
    """Per-variable weight tables for ssg_plan_set_filter: (fourier_w [nvars][K + 1] or
    None for an empty fourier_filter, chebyshev_w [nvars][b_kDim] or None); 1.0 for a
    variable the dict does not cover.  names: the variables in index order."""
    def table(filters, n):
        if not filters:
            return None
        w = np.ones((len(names), n), dtype=np.float64)
        for v, name in enumerate(names):
            f = get_filter(filters, name)
            if f is not None:
                w[v] = [filter_weight(f, k) for k in range(n)]
        return w
    return table(gp.fourier_filter, K + 1), table(gp.chebyshev_filter, b_kDim)
