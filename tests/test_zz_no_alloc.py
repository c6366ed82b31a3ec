"""GPU: ssg_grid_transform_noalloc, the allocation-free variant of
ssg_grid_transform (intermediates in the library's static scratch array).

Through the C-ABI, its output must equal ssg_grid_transform's bitwise (same
kernels, same launch shapes) on two calls in a row, the second on a smaller
grid, so that the second call runs on an array that still holds the first
call's larger contents; an RZ grid and an R3X grid with background
coefficients follow; FP32 plans are refused with SSG_EINVAL, and a plan whose
intermediates exceed the array with SSG_ENOTSUP.
"""
import numpy as np
import pytest

import springsteel_hip as S
from springsteel_hip.errors import NotSupportedError
from parity_util import seeded_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _arrays(grid, seed):
    nv = grid.resolved.nvars
    phys = torch.zeros((grid.nslots, nv, grid.npts), dtype=torch.float64, device="cuda")
    spec = torch.zeros((nv, grid.nspec), dtype=torch.float64, device="cuda")
    f = seeded_field(S.getGridpoints(grid), nv, seed, grid.geometry)
    phys[0].copy_(torch.from_numpy(np.ascontiguousarray(f.T)))
    S.spectralTransform_device(grid, phys, spec)
    torch.cuda.synchronize()
    return phys, spec


def _compare(grid, seed):
    phys, spec = _arrays(grid, seed)
    ref = torch.zeros_like(phys)
    grid.plan.grid(spec.data_ptr(), ref.data_ptr(), S.SSG_MEM_DEVICE, torch.cuda.current_stream().cuda_stream)
    out = torch.full_like(phys, float("nan"))
    grid.plan.grid_noalloc(spec.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    same = torch.equal(out, ref)
    print(f"\n{grid.geometry} nc={grid.params.num_cells} vars={grid.resolved.nvars} precision={grid.precision}: "
          f"noalloc {'==' if same else '!='} ssg_grid_transform")
    assert same


def test_noalloc_matches_grid_transform_twice(cuda):
    precision = S.SSG_FP64
    big = S.createGrid(S.SpringsteelGridParameters(geometry="RLZ", num_cells=14, iMin=0.0, iMax=70.0, kMin=0.0,
                                                   kMax=10.0, kDim=16, vars={"u": 1, "v": 2}), precision=precision)
    small = S.createGrid(S.SpringsteelGridParameters(geometry="RL", num_cells=6, iMin=0.0, iMax=30.0,
                                                     vars={"u": 1}), precision=precision)
    _compare(big, 3)
    _compare(small, 4)          # the array still holds the larger grid's contents
    _compare(big, 5)
    big.close()
    small.close()


def test_noalloc_rz_and_r3x_ahat(cuda):
    rz = S.createGrid(S.SpringsteelGridParameters(geometry="RZ", num_cells=9, iMin=0.0, iMax=45.0, kMin=0.0,
                                                  kMax=12.0, kDim=13, vars={"u": 1, "v": 2}))
    _compare(rz, 6)
    rz.close()
    CB = S.CubicBSpline
    g = S.createGrid(S.SpringsteelGridParameters(geometry="RL", num_cells=8, iMin=0.0, iMax=40.0, vars={"u": 1},
                                                 BCL={"u": CB.R3X}, BCR={"u": CB.R0}))
    rng = np.random.default_rng(2)
    S.set_ahat(g, np.asfortranarray(rng.normal(size=(g.nspec, 1))))
    _compare(g, 7)               # btilde in the scratch as well
    g.close()


def test_noalloc_refuses_fp32_plans(cuda):
    # FP64 plans only: the FP32 form's physical arrays are Float32
    from springsteel_hip.errors import ArgumentError
    g = S.createGrid(S.SpringsteelGridParameters(geometry="RL", num_cells=6, iMin=0.0, iMax=30.0, vars={"u": 1}),
                     precision=S.SSG_FP32_REFINE)
    phys = torch.zeros((g.nslots, 1, g.npts), dtype=torch.float32, device="cuda")
    spec = torch.zeros((1, g.nspec), dtype=torch.float64, device="cuda")
    with pytest.raises(ArgumentError):
        g.plan.grid_noalloc(spec.data_ptr(), phys.data_ptr(), torch.cuda.current_stream().cuda_stream)
    g.close()


def test_noalloc_refuses_plans_beyond_the_array(cuda):
    g = S.createGrid(S.SpringsteelGridParameters(geometry="RLZ", num_cells=120, iMin=0.0, iMax=500.0, kMin=0.0,
                                                 kMax=10.0, kDim=32, vars={"u": 1}))
    nv = g.resolved.nvars
    phys = torch.zeros((g.nslots, nv, g.npts), dtype=torch.float64, device="cuda")
    spec = torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda")
    with pytest.raises(NotSupportedError, match="static"):
        g.plan.grid_noalloc(spec.data_ptr(), phys.data_ptr(), torch.cuda.current_stream().cuda_stream)
    g.close()
