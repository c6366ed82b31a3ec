"""GPU parity: the HIP library against the CPU oracle, FP64 plain form.

Bar: relative L2 <= 1e-12 for spectral (per variable) and for every physical
slot (per variable), on the reference's own test grids (test/transforms.jl
380-907), on BC variants, and on the RL / RZ workloads at full size.
"""
import numpy as np
import pytest

import springsteel_hip as S
from springsteel_hip import workloads as W
from parity_util import TOL_FP64, errors, run_both, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu

CB = S.CubicBSpline
CH = S.Chebyshev


def _check(gp, field=None, seed=0, tol=TOL_FP64):
    grid, og, spec, ospec, phys, ophys, phys2, inp = run_both(gp, field, seed)
    es = errors(spec, ospec)
    ep = errors(phys, ophys, grid.geometry, grid.params)
    ep2 = errors(phys2, ophys, grid.geometry, grid.params)
    sens = sensitivity(og, inp, ophys, grid.geometry, grid.params)
    tols = slot_tolerances(sens, tol)
    ratio = max((max(ep[v][s], ep2[v][s]) / sens[v][s] for v in range(len(ep)) for s in range(phys.shape[2])
                 if sens[v][s] > 0 and max(ep[v][s], ep2[v][s]) > tol), default=0.0)
    print(f"\nworst error / rounding sensitivity over slots above {tol:.0e}: {ratio:.2f}")
    print(f"\n{gp.geometry} nc={gp.num_cells} kDim={gp.kDim} nvars={len(gp.vars)}: spectral {max(es):.2e}, "
          f"slots {[f'{max(e[s] for e in ep):.1e}' for s in range(phys.shape[2])]}, "
          f"tolerances {[f'{max(t[s] for t in tols):.1e}' for s in range(phys.shape[2])]}")
    assert max(es) <= tol, es
    for v in range(len(ep)):
        for s in range(phys.shape[2]):
            assert ep[v][s] <= tols[v][s], (v, s, ep[v], tols[v])
            assert ep2[v][s] <= tols[v][s], (v, s, ep2[v], tols[v])
    grid.close()


# reference test grids (test/transforms.jl)
def test_rl_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1}))


def test_rz_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RZ", num_cells=8, iMin=0.0, iMax=40.0, kMin=0.0, kMax=20.0,
                                       kDim=15, vars={"u": 1}))


def test_rz_multivar_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RZ", num_cells=6, iMin=0.0, iMax=30.0, kMin=0.0, kMax=15.0,
                                       kDim=12, vars={"u": 1, "v": 2}))


def test_rlz_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                       kDim=8, vars={"u": 1}))


def test_rlz_patch_offset():
    # spectralIndexL = 2 -> patchOffsetL = 3 (test/transforms.jl:826-869)
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                       kDim=8, spectralIndexL=2, vars={"u": 1}))


def test_rl_patch_offset_and_cap():
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=12, iMin=10.0, iMax=100.0, spectralIndexL=3,
                                       max_wavenumber={"u": 5, "w": -1}, vars={"u": 1, "w": 2}))


def test_rlz_wavenumber_cap():
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=6, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                       kDim=9, max_wavenumber={"default": 4}, vars={"u": 1, "v": 2}))


@pytest.mark.parametrize("bcl,bcr", [
    (CB.R1T0, CB.R1T1), (CB.R1T1, CB.R0), (CB.R2T10, CB.R1T2), (CB.R0, CB.R2T20), (CB.R3, CB.R1T0),
])
def test_rl_spline_bcs(bcl, bcr):
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=9, iMin=0.0, iMax=90.0, vars={"u": 1},
                                       BCL={"u": bcl}, BCR={"u": bcr}))


@pytest.mark.parametrize("bcb,bct", [
    (CH.R0, CH.R0), (CH.R1T0, CH.R0), (CH.R1T1, CH.R0), (CH.R0, CH.R1T0), (CH.R0, CH.R1T1),
    (CH.R1T0, CH.R1T0), (CH.R1T1, CH.R1T1), (CH.R1T0, CH.R1T1), (CH.R1T1, CH.R1T0),
])
def test_rz_chebyshev_bcs(bcb, bct):
    _check(S.SpringsteelGridParameters(geometry="RZ", num_cells=7, iMin=0.0, iMax=35.0, kMin=0.0, kMax=12.0,
                                       kDim=13, vars={"u": 1}, BCB={"u": bcb}, BCT={"u": bct}))


def test_rlz_mixed_bcs_multivar():
    _check(S.SpringsteelGridParameters(
        geometry="RLZ", num_cells=5, iMin=0.0, iMax=50.0, kMin=0.0, kMax=10.0, kDim=10,
        vars={"u": 1, "v": 2, "w": 3}, BCL={"u": CB.R1T1, "default": CB.R0}, BCR={"v": CB.R1T0, "default": CB.R0},
        BCB={"w": CH.R1T1, "default": CH.R0}, BCT={"u": CH.R1T0, "default": CH.R0}, l_q={"v": 3.0, "default": 2.0}))


@pytest.mark.parametrize("kdim", [48, 64])
def test_rlz_many_heights(kdim):
    # more than 32 heights: the forward Chebyshev stage's 32-point tiles (5 or 4
    # points per load pass) and the inverse stage's largest matrices
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=40.0, kMin=0.0, kMax=4.0,
                                       kDim=kdim, vars={"u": 1, "v": 2}))


def test_regular_quadrature_rl():
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=8, iMin=0.0, iMax=80.0, mubar=2,
                                       quadrature="regular", vars={"u": 1}))


@pytest.mark.parametrize("geom", ["RL", "RLZ"])
def test_regular_quadrature_six_points(geom):
    # :regular rules may carry more points per cell than the Gauss tables (test/mubar.jl:34-35)
    kw = dict(geometry=geom, num_cells=6, iMin=0.0, iMax=30.0, mubar=6, quadrature="regular", vars={"u": 1})
    if geom == "RLZ":
        kw.update(kMin=0.0, kMax=5.0, kDim=8)
    _check(S.SpringsteelGridParameters(**kw))


def test_rl_vortex_full():
    gp = W.rl_params()
    _check(gp, field=lambda pts, nv: W.rl_field(pts, nv))


def test_rz_modes_full():
    gp = W.rz_params()
    _check(gp, field=lambda pts, nv: W.rz_field(pts, nv, L=gp.iMax, zmax=gp.kMax))


def test_rlz_reduced_workload_field():
    # the workload field is linear in z: slot 7 is measured against its scale floor
    gp = W.rlz_params(num_cells=40, nvars=2)
    _check(gp, field=lambda pts, nv: W.rlz_field(pts, nv))


def test_rlz_reduced_curved_field():
    # seeded field with curvature in r, lambda and z: every slot carries signal
    _check(W.rlz_params(num_cells=30, nvars=2), seed=5)
