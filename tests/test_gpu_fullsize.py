"""GPU parity at the full sizes of the three RLZ workloads.

* The RLZ workload (200 cells on [0, 500], rings 4 + 4r up to 2404 points, 32
  Chebyshev points, 5 variables, 23,155,200 points per variable): every
  variable and every one of the 7 slots against the oracle at the FP64 bar
  (1e-12 relative L2, or the slot's rounding-sensitivity tolerance where the
  derivative operator amplifies rounding beyond it), spectral per variable at
  1e-12.
* The FP32 form on the same grid and input (values representable in Float32,
  the form's storage): per variable and per derivative order (d0 = slot 1,
  d1 = slots 2/4/6, d2 = slots 3/5/7) against the oracle, printed and held to
  1e-6; spectral (the form's FP64 forward path) to 1e-12.
* The 64-grid batch (workloads.batch_params: num_cells ~ U{50..400} scaled to
  fit one GPU, nz in {1, 16, 32, 64}, 1-6 variables), both forms: every grid's
  spectral and physical arrays bitwise equal to its own single-grid calls, and
  the smallest RL and RLZ grids against the oracle at their form's bar.

The oracle runs at up to 16 threads; the full-size oracle round trips take
tens of seconds each, so these tests carry their own time limits.
"""
import os

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip import capi, workloads as W
from springsteel_hip.metrics import rel_l2_slots
from parity_util import TOL_FP64, sensitivity, slot_tolerances

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

THREADS = min(len(os.sched_getaffinity(0)), 16)
TOL_FP32 = 1e-6
ORDERS = {"d0": (0,), "d1": (1, 3, 5), "d2": (2, 4, 6)}


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def rlz_full(cuda):
    gp = W.rlz_params()
    grid = S.createGrid(gp)
    # input values representable in Float32, so that both forms (the FP32 form stores
    # physical in Float32) transform the same numbers and one oracle run serves both
    field = np.asfortranarray(W.workload_input("rlz_vortex", grid).astype(np.float32).astype(np.float64))
    og = O.OracleGrid(grid.resolved)
    inp = og.new_physical()
    inp[:, :, 0] = field
    _log(f"\n[fullsize] RLZ workload: oracle round trip, {grid.resolved.nvars} variables, {THREADS} threads")
    ospec = og.spectral_transform(inp, threads=THREADS)
    ophys = og.grid_transform(ospec, threads=THREADS)
    _log("[fullsize] oracle rounding sensitivity (perturbed round trip)")
    sens = sensitivity(og, inp, ophys, grid.geometry, grid.params)
    del inp
    yield dict(params=gp, grid=grid, field=field, ospec=ospec, ophys=ophys, sens=sens)
    grid.close()


def _dtype(g):
    return torch.float32 if g.precision == S.SSG_FP32_REFINE else torch.float64


def _gpu_round_trip(grid, field):
    nv = grid.resolved.nvars
    phys = torch.zeros((grid.nslots, nv, grid.npts), dtype=_dtype(grid), device="cuda")
    spec = torch.zeros((nv, grid.nspec), dtype=torch.float64, device="cuda")
    phys[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
    S.spectralTransform_device(grid, phys, spec)
    S.gridTransform_device(grid, spec, phys)
    torch.cuda.synchronize()
    return phys, spec


def _errors(grid, phys, spec, ospec, ophys):
    """per variable: spectral rel L2 and the floored rel L2 of every slot"""
    out = []
    for v in range(grid.resolved.nvars):
        gs = spec[v].cpu().numpy()
        es = float(np.linalg.norm(gs - ospec[:, v]) / np.linalg.norm(ospec[:, v]))
        gp = phys[:, v, :].double().cpu().numpy().T
        out.append((es, rel_l2_slots(gp, ophys[:, v, :], grid.geometry, grid.params)))
    return out


@pytest.mark.timeout(900)
def test_rlz_workload_full_fp64(rlz_full):
    d = rlz_full
    grid = d["grid"]
    _log("[fullsize] RLZ workload: FP64 form on the GPU")
    phys, spec = _gpu_round_trip(grid, d["field"])
    errs = _errors(grid, phys, spec, d["ospec"], d["ophys"])
    tols = slot_tolerances(d["sens"])
    for v, (es, ep) in enumerate(errs):
        _log(f"  var {v + 1}: spectral {es:.2e}, slots {[f'{e:.1e}' for e in ep]}, "
             f"tolerances {[f'{t:.1e}' for t in tols[v]]}")
    for v, (es, ep) in enumerate(errs):
        assert es <= TOL_FP64, (v, es)
        for s in range(grid.nslots):
            assert ep[s] <= tols[v][s], (v, s, ep, tols[v])


@pytest.mark.timeout(900)
def test_rlz_workload_full_fp32(rlz_full):
    d = rlz_full
    g32 = S.createGrid(d["params"], precision=S.SSG_FP32_REFINE)
    _log("[fullsize] RLZ workload: FP32 form on the GPU")
    phys, spec = _gpu_round_trip(g32, d["field"])
    errs = _errors(g32, phys, spec, d["ospec"], d["ophys"])
    for v, (es, ep) in enumerate(errs):
        by_order = {o: max(ep[s] for s in ss) for o, ss in ORDERS.items()}
        _log(f"  var {v + 1}: spectral {es:.2e}, " + ", ".join(f"{o} {e:.2e}" for o, e in by_order.items()))
    for v, (es, ep) in enumerate(errs):
        assert es <= TOL_FP64, (v, es)
        for s in range(g32.nslots):
            assert ep[s] <= TOL_FP32, (v, s, ep[s])
    g32.close()


def _batch_inputs(params, seeds):
    """rl_field on each grid's ring points, constant in z, as device tensors (nvars, npts) of
    values representable in Float32 (the FP32 form's storage)"""
    out = []
    for gp, sd in zip(params, seeds):
        rl = S.SpringsteelGridParameters(**{**gp.__dict__, "geometry": "RL", "jDim": 0, "b_jDim": 0,
                                            "kDim": 0, "b_kDim": 0})
        pts = S.getGridpoints(S.SpringsteelGrid(S.compute_derived_params(rl)))
        base = W.rl_field(pts, len(gp.vars), seed=sd).astype(np.float32).astype(np.float64)
        b = torch.from_numpy(np.ascontiguousarray(base.T)).cuda()
        kz = gp.kDim if gp.geometry == "RLZ" else 1
        out.append(b.unsqueeze(2).expand(b.shape[0], b.shape[1], kz).reshape(b.shape[0], -1).contiguous())
    return out


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("precision", [S.SSG_FP64, S.SSG_FP32_REFINE], ids=["fp64", "fp32"])
def test_batch64_full(cuda, precision):
    params, seeds, total, scale = W.batch_params()
    _log(f"\n[fullsize] batch of {len(params)} grids, num_cells scale {scale:.3f}, "
         f"FP64 footprint {total / 1e9:.0f} GB, form {precision}")
    grids = [S.createGrid(gp, precision=precision) for gp in params]
    inputs = [x.to(_dtype(g)) for x, g in zip(_batch_inputs(params, seeds), grids)]
    phys, spec = [], []
    for g in grids:
        nv = g.resolved.nvars
        phys.append(torch.zeros((g.nslots, nv, g.npts), dtype=_dtype(g), device="cuda"))
        spec.append(torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda"))
    batch = capi.Batch(grids, phys, spec)
    for ph, x in zip(phys, inputs):
        ph[0].copy_(x)
    torch.cuda.synchronize()
    batch.spectral()
    batch.grid()
    torch.cuda.synchronize()
    _log("[fullsize] batch done; single-grid calls")
    # every grid: bitwise its own single-grid calls on the same arrays
    for i, g in enumerate(grids):
        bsp, bph = spec[i].clone(), phys[i].clone()
        phys[i].zero_()
        spec[i].zero_()
        phys[i][0].copy_(inputs[i])
        S.spectralTransform_device(g, phys[i], spec[i])
        S.gridTransform_device(g, spec[i], phys[i])
        torch.cuda.synchronize()
        assert torch.equal(bsp, spec[i]), f"grid {i}: spectral differs from its single-grid call"
        assert torch.equal(bph, phys[i]), f"grid {i}: physical differs from its single-grid call"
        del bsp, bph
    batch.close()
    # the smallest RL and RLZ grids against the oracle
    for geo in ("RL", "RLZ"):
        i = min((k for k, g in enumerate(grids) if g.geometry == geo), key=lambda k: grids[k].npts)
        g = grids[i]
        og = O.OracleGrid(g.resolved)
        hp = og.new_physical()
        hp[:, :, 0] = inputs[i].double().cpu().numpy().T
        ospec = og.spectral_transform(hp, threads=THREADS)
        ophys = og.grid_transform(ospec, threads=THREADS)
        sens = sensitivity(og, hp, ophys, g.geometry, g.params)
        errs = _errors(g, phys[i], spec[i], ospec, ophys)
        _log(f"  grid {i} ({geo}, nc {g.params.num_cells}, kDim {g.params.kDim}, {g.resolved.nvars} vars): "
             f"spectral {max(e[0] for e in errs):.2e}, slots "
             f"{[f'{max(e[1][s] for e in errs):.1e}' for s in range(g.nslots)]}")
        for v, (es, ep) in enumerate(errs):
            if precision == S.SSG_FP64:
                tols = slot_tolerances(sens)
                assert es <= TOL_FP64, (i, v, es)
                for s in range(g.nslots):
                    assert ep[s] <= tols[v][s], (i, v, s, ep, tols[v])
            else:
                assert es <= TOL_FP64, (i, v, es)
                for s in range(g.nslots):
                    assert ep[s] <= TOL_FP32, (i, v, s, ep[s])
    for g in grids:
        g.close()
    del phys, spec, inputs
    torch.cuda.empty_cache()
