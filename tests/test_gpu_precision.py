"""GPU: the FP32 form (SSG_FP32_REFINE) of the ring grids against the FP64 oracle.

The form stores physical in Float32 (input slot and the output slots); its
forward path (forward Chebyshev, ring DFT, projection), the banded solve and
the radial evaluation run the FP64 kernels, and the ring and Chebyshev
synthesis runs in FP32 on a Float32 copy of the evaluated values.  The oracle
runs on the same input values (the Float32 input, widened), so the errors are
those of the form's arithmetic and output rounding alone.

Bars, relative L2 with the slot scale floors:
  spectral                      <= 1e-12 per variable (FP64 path);
  every physical slot (d0/d1/d2) <= 1e-6 per variable.
"""
import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import capi, workloads as W
from springsteel_hip.errors import ArgumentError
from springsteel_hip.metrics import rel_l2_slots
from parity_util import TOL_FP64, seeded_field

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-6


def _run(gp, field=None, seed=0):
    g32 = S.createGrid(gp, precision=S.SSG_FP32_REFINE)
    og = O.OracleGrid(g32.resolved)
    pts = S.getGridpoints(g32)
    nv = g32.resolved.nvars
    f = seeded_field(pts, nv, seed, g32.geometry) if field is None else field(pts, nv)
    g32.physical[:, :, 0] = f                          # rounded to float32 here
    inp = og.new_physical()
    inp[:, :, 0] = g32.physical[:, :, 0]               # the same values, widened
    S.spectralTransform_(g32)
    S.gridTransform_(g32)
    ospec = og.spectral_transform(inp, threads=8)
    ophys = og.grid_transform(ospec, threads=8)
    gphys = g32.physical.astype(np.float64)
    errs = [rel_l2_slots(gphys[:, v, :], ophys[:, v, :], g32.geometry, g32.params) for v in range(nv)]
    spec_err = [float(np.linalg.norm(g32.spectral[:, v] - ospec[:, v]) / np.linalg.norm(ospec[:, v]))
                for v in range(nv)]
    g32.close()
    return spec_err, errs


def _check(gp, field=None, seed=0):
    spec_err, errs = _run(gp, field, seed)
    print(f"\nFP32 form {gp.geometry} nc={gp.num_cells}: spectral {max(spec_err):.2e}, slots "
          f"{[f'{max(e[s] for e in errs):.1e}' for s in range(len(errs[0]))]}")
    assert max(spec_err) <= TOL_FP64, spec_err
    for v in range(len(errs)):
        for s in range(len(errs[v])):
            assert errs[v][s] <= TOL_FP32, (v, s, errs[v][s])


def test_fp32_rl_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1}))


def test_fp32_rl_multivar_bcs_cap():
    CB = S.CubicBSpline
    _check(S.SpringsteelGridParameters(geometry="RL", num_cells=12, iMin=0.0, iMax=60.0,
                                       vars={"u": 1, "v": 2, "w": 3}, BCL={"default": CB.R0, "v": CB.R1T0},
                                       BCR={"default": CB.R0, "w": CB.R1T1}, max_wavenumber={"default": -1, "u": 4}))


def test_fp32_rlz_reference_grid():
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=4, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0,
                                       kDim=8, vars={"u": 1}))


def test_fp32_rlz_reduced_workload_field():
    gp = W.rlz_params(num_cells=40, nvars=2)
    _check(gp, field=lambda pts, nv: W.rlz_field(pts, nv))


def test_fp32_rlz_curved_field_and_folds():
    CH = S.Chebyshev
    _check(S.SpringsteelGridParameters(geometry="RLZ", num_cells=24, iMin=0.0, iMax=120.0, kMin=0.0, kMax=50.0,
                                       kDim=20, vars={"u": 1, "v": 2},
                                       BCB={"default": CH.R0, "v": CH.R1T1}, BCT={"default": CH.R0, "u": CH.R1T0}),
           seed=7)


def test_tuning_set_get_and_range_check():
    g = S.createGrid(S.SpringsteelGridParameters(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1}),
                     precision=S.SSG_FP32_REFINE)
    assert capi.plan_tuning(g.plan) == {k: 0 for k in capi.TUNING_KEYS}
    capi.plan_set_tuning(g.plan, syn_ct=2, syn_nwq=4, ana_mr=1, si_seg=2, sa_mode=2)
    assert capi.plan_tuning(g.plan) == {"syn_ct": 2, "syn_nwq": 4, "ana_mr": 1, "si_seg": 2, "sa_mode": 2}
    for bad in ({"syn_nwq": 3}, {"si_seg": 3}, {"sa_mode": 1}, {"sa_mode": 3}, {"syn_ct": -1}):
        with pytest.raises(ArgumentError):
            capi.plan_set_tuning(g.plan, **bad)
    assert capi.plan_tuning(g.plan)["si_seg"] == 2        # a rejected call changes nothing
    g.close()


def test_forms_refuse_the_other_element_type():
    torch = pytest.importorskip("torch")
    gp = S.SpringsteelGridParameters(geometry="RL", num_cells=6, iMin=0.0, iMax=30.0, vars={"u": 1})
    g32 = S.createGrid(gp, precision=S.SSG_FP32_REFINE)
    g64 = S.createGrid(gp)
    ph32 = torch.zeros((g32.nslots, 1, g32.npts), dtype=torch.float32, device="cuda")
    ph64 = torch.zeros((g64.nslots, 1, g64.npts), dtype=torch.float64, device="cuda")
    sp = torch.zeros((1, g64.nspec), dtype=torch.float64, device="cuda")
    with pytest.raises(ArgumentError):
        S.spectralTransform_device(g32, ph64, sp)
    with pytest.raises(ArgumentError):
        S.gridTransform_device(g64, sp, ph32)
    # at the C-ABI: the FP64 entry points refuse an FP32 plan and the reverse
    lib = capi.load()
    st = torch.cuda.current_stream().cuda_stream
    import ctypes
    rc = lib.ssg_grid_transform(g32.plan.handle, ctypes.c_void_p(sp.data_ptr()), ctypes.c_void_p(ph32.data_ptr()),
                                capi.SSG_MEM_DEVICE, ctypes.c_void_p(st))
    assert rc == capi.SSG_EINVAL
    rc = lib.ssg_spectral_transform_f32(g64.plan.handle, ctypes.c_void_p(ph64.data_ptr()),
                                        ctypes.c_void_p(sp.data_ptr()), capi.SSG_MEM_DEVICE, ctypes.c_void_p(st))
    assert rc == capi.SSG_EINVAL
    g32.close()
    g64.close()


def test_fp32_rl_vortex_full():
    gp = W.rl_params()
    _check(gp, field=lambda pts, nv: W.rl_field(pts, nv))
