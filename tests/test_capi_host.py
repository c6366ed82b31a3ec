"""Host-side tests of the C-ABI library and the Python grid mirror (no GPU).

The library must load and export every entry point include/springsteel_hip.h
declares.  Descriptor validation runs on the host before any HIP call, so the
reference's failure modes (DomainError, ArgumentError, BoundsError, and the
features this library rejects with ENOTSUP) are injected here through the
C-ABI itself.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import capi
from springsteel_hip.params import compute_derived_params, resolve

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CB, CH = S.CubicBSpline, S.Chebyshev


def header_functions():
    text = open(os.path.join(REPO, "include", "springsteel_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ssg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = header_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.EXPORTED)


def _header_arity():
    """Number of parameters of every declared entry point."""
    text = open(os.path.join(REPO, "include", "springsteel_hip.h")).read()
    out = {}
    for m in re.finditer(r"^\s*(?:int|const char\*)\s+(ssg_\w+)\s*\(([^)]*)\)", text, flags=re.M):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def _julia_ccalls():
    """(symbol, number of argument types) of every ccall in the Julia binding."""
    text = open(os.path.join(REPO, "springsteel.jl_amd", "julia", "SpringsteelHIP.jl")).read()
    out = []
    for m in re.finditer(r"ccall\(\(:(ssg_\w+), libssg\),\s*\w+,\s*\(", text):
        i, depth, j = m.end(), 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(text[j], 0)
            j += 1
        types = text[i:j - 1].strip().rstrip(",")
        # the types hold no commas of their own (Ptr{Ptr{Cvoid}}, Ref{SsgDesc}); a
        # one-element tuple is written (T,)
        out.append((m.group(1), 0 if not types else types.count(",") + 1))
    return out


def test_julia_binding_matches_header():
    arity = _header_arity()
    calls = _julia_ccalls()
    assert len(calls) >= 20
    for name, n in calls:
        assert name in arity, name
        assert n == arity[name], (name, n, arity[name])
    # every entry point of the tile pipeline is bound
    bound = {n for n, _ in calls}
    for name in ("ssg_spline_transform", "ssg_tile_transform", "ssg_sum_spectral_tile", "ssg_set_spectral_tile",
                 "ssg_sum_shared_spectral"):
        assert name in bound, name


def _create(gp):
    rg = resolve(compute_derived_params(gp))
    return capi.Plan(rg)


@pytest.mark.parametrize("gp,exc", [
    (S.SpringsteelGridParameters(geometry="RL", num_cells=4, iMin=0.0, iMax=1.0, BCL={"default": CB.PERIODIC}),
     S.NotSupportedError),
    (S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=1.0, kMin=0.0, kMax=1.0, kDim=3),
     S.BoundsError),
    (S.SpringsteelGridParameters(geometry="RZ", num_cells=4, iMin=0.0, iMax=1.0, kMin=0.0, kMax=1.0, kDim=1),
     S.ArgumentError),
    (S.SpringsteelGridParameters(geometry="RZ", num_cells=4, iMin=0.0, iMax=1.0, kMin=0.0, kMax=1.0, kDim=8,
                                 BCB={"default": CH.R1T2}), S.DomainError),
    (S.SpringsteelGridParameters(geometry="RL", num_cells=4, iMin=0.0, iMax=1.0, mubar=6), S.ArgumentError),
    (S.SpringsteelGridParameters(geometry="RL", num_cells=4, iMin=1.0, iMax=1.0), S.ArgumentError),
])
def test_plan_create_rejects(gp, exc):
    with pytest.raises(exc):
        _create(gp)


def test_filter_flags_in_the_descriptor_rejected():
    """Filter weights go through ssg_plan_set_filter; a descriptor asking for filters is refused."""
    rg = resolve(compute_derived_params(S.SpringsteelGridParameters(geometry="RL", num_cells=4, iMin=0.0, iMax=1.0)))
    d, keep = capi.make_desc(rg)
    d.filter_flags = 1
    h = ctypes.c_void_p()
    assert capi.load().ssg_plan_create(ctypes.byref(d), ctypes.byref(h)) == capi.SSG_ENOTSUP
    # a value that is not a filter fails in the host weights, as _filter_weight has no method for it
    from springsteel_hip.filters import filter_weights
    gp = compute_derived_params(S.SpringsteelGridParameters(geometry="RL", num_cells=4, iMin=0.0, iMax=1.0,
                                                            fourier_filter={"u": 1}))
    with pytest.raises(S.ArgumentError):
        filter_weights(gp, ["u"], gp.iDim, gp.b_kDim)


def test_fp32_form_rejected_on_rz_and_unknown_precision():
    rz = S.SpringsteelGridParameters(geometry="RZ", num_cells=4, iMin=0.0, iMax=1.0, kMin=0.0, kMax=1.0, kDim=8)
    rg = resolve(S.compute_derived_params(rz))
    with pytest.raises(S.NotSupportedError):
        capi.Plan(rg, capi.SSG_FP32_REFINE)
    with pytest.raises(S.ArgumentError):
        capi.Plan(rg, 7)


def test_unknown_geometry_is_domain_error():
    with pytest.raises(S.DomainError):
        resolve(S.SpringsteelGridParameters(geometry="RRR", num_cells=4, iMin=0.0, iMax=1.0))


def test_rlz_differing_wavenumber_caps_rejected():
    gp = S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=1.0, kDim=8, kMax=1.0,
                                     vars={"u": 1, "v": 2}, max_wavenumber={"u": 3, "v": 4})
    with pytest.raises(S.NotSupportedError):
        resolve(gp)


def test_bc_conversion_table():
    # src/factory.jl:343-396
    k = S.bcs.spline_bc_kind
    assert k(S.NaturalBC()) == (0, 0.0, 0.0)
    assert k(S.DirichletBC()) == (1, -4.0, -1.0)
    assert k(S.NeumannBC()) == (1, 0.0, 1.0)
    assert k(S.SecondDerivativeBC()) == (1, 2.0, -1.0)
    assert k(S.CauchyBC(0.0, 0.0)) == (2, 1.0, -0.5)
    assert k(S.AntisymmetricBC()) == (2, -1.0, 0.0)
    assert k(S.ZerosBC()) == (4, 0.0, 0.0)
    assert k(S.FixedBC()) == (5, 0.0, 0.0)
    lam = 2.0
    kind, a, b = k(S.ExponentialBC(lam))
    assert kind == 1 and abs(a + 4.0 / (3 * lam + 1)) < 1e-15 and abs(b - (3 * lam - 1) / (3 * lam + 1)) < 1e-15
    assert k(CB.R1T0) == (1, -4.0, -1.0) and k(CB.R2T20) == (2, -1.0, 0.0) and k(CB.R3) == (4, 0.0, 0.0)
    c = S.bcs.chebyshev_bc_kind
    assert c(S.NaturalBC()) == 0 and c(S.DirichletBC(1.0)) == 1 and c(S.NeumannBC()) == 2
    assert c(CH.R1T2) == 3
    with pytest.raises(S.ArgumentError):
        c(S.PeriodicBC())
    with pytest.raises(S.ArgumentError):
        S.bcs.spline_bc_kind(S.BoundaryConditions(du=0.0, d2u=0.0))


def test_per_variable_resolution_with_default():
    gp = S.SpringsteelGridParameters(geometry="RZ", num_cells=5, iMin=0.0, iMax=1.0, kDim=6, kMax=1.0,
                                     vars={"a": 2, "b": 1}, l_q={"a": 3.0, "default": 2.5},
                                     BCL={"b": CB.R1T1, "default": CB.R0}, BCT={"a": CH.R1T0, "default": CH.R0})
    rg = resolve(gp)
    assert rg.nvars == 2
    assert rg.l_q == [2.5, 3.0]                       # var 1 = "b", var 2 = "a"
    assert rg.bcl[0] == (1, 0.0, 1.0) and rg.bcl[1] == (0, 0.0, 0.0)
    assert rg.bct == [0, 1]


@pytest.mark.parametrize("geom,kw", [("RL", {}), ("RZ", {"kDim": 9, "kMax": 5.0}), ("RLZ", {"kDim": 6, "kMax": 5.0}),
                                     ("RLZ", {"kDim": 6, "kMax": 5.0, "spectralIndexL": 2})])
def test_gridpoints_and_dims_match_oracle(geom, kw):
    gp = compute_derived_params(S.SpringsteelGridParameters(geometry=geom, num_cells=5, iMin=1.0, iMax=11.0, **kw))
    g = S.SpringsteelGrid(gp)
    og = O.OracleGrid(g.resolved)
    assert (g.npts, g.nspec, g.nslots) == (og.npts, og.nspec, og.nslots)
    assert np.allclose(S.getGridpoints(g), og.gridpoints(), rtol=0, atol=1e-12)


def test_derived_dims_match_reference_formulas():
    # src/factory.jl:166-175 and the RLZ config: jDim 723,600, 23,155,200 points per variable
    from springsteel_hip import workloads
    gp = compute_derived_params(workloads.rlz_params())
    assert gp.iDim == 600 and gp.b_iDim == 203 and gp.b_kDim == 22
    assert gp.jDim == 723600 and gp.b_jDim == sum(2 * r + 1 for r in range(1, 601))
    g = S.SpringsteelGrid(gp)
    assert g.npts == 23155200 and g.nspec == 22 * 203 * 1201
    gz = S.SpringsteelGrid(compute_derived_params(workloads.rz_params()))
    assert gz.params.b_kDim == 43 and gz.nspec == 43 * 259
