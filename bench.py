#!/usr/bin/env python3
"""Throughput of the RLZ round trip (spectralTransform! + gridTransform!) on one MI355X.

Workload (default): RLZ, 200 cells on [0, 500] (600 radial points, rings 4+4r up
to 2404 points), 32 Chebyshev points on [0, 20000], 5 variables; seeded Rankine-
like vortex plus azimuthal modes k <= 16 varying linearly in z.  One step = one
device-resident round trip over all variables: 115,776,000 point-variables in,
7 slots out.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rlz_vortex|rz_modes|rl_vortex|slz_sphere]
                  [--no-cpu] [--parity]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "springsteel.jl_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X dense FP64 (vector and matrix)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(grid):
    """Algorithmic HBM bytes per launch of each stage (8-byte values): the stage's
    compulsory input + output arrays."""
    from springsteel_hip.params import ring_layout
    p = grid.params
    nv, bk = grid.resolved.nvars, max(p.b_kDim, 1)
    geom = {"SL": "RL", "SLZ": "RLZ"}.get(grid.geometry, grid.geometry)
    nzb = 1 if geom == "RL" else bk
    ntr = nv * nzb
    phys1 = grid.npts * nv * 8
    spec = grid.nspec * nv * 8
    out = {}
    if geom == "RZ":
        cb = bk * p.iDim * nv * 8
        out["cheb_fwd"] = phys1 + cb
        out["sb"] = cb + spec
        out["sa_si"] = spec + 3 * cb
        out["cheb_inv"] = 3 * cb + grid.nslots * phys1
    else:
        rc = ntr * p.b_jDim * 8
        # synthesis input: (P, Q) of 3 orders per (ring, k <= ri, transform)
        xs = 6 * ntr * int(sum(kr + 1 for _, kr, _ in ring_layout(p))) * 8
        if geom == "RLZ":
            cm = nv * p.jDim * bk * 8
            out["cheb_fwd"] = phys1 + cm
            out["ring_ana"] = cm + rc
        else:
            out["ring_ana"] = phys1 + rc
        out["sb"] = rc + spec
        out["sa_si"] = spec + xs
        out["ring_syn_cheb" if geom == "RLZ" else "ring_syn"] = xs + grid.nslots * phys1
        if geom == "RLZ":
            # chirp-z synthesis: X in, the five families of every (ring point, mode) out (Y);
            # the inverse Chebyshev stage reads Y and writes every slot
            ys = 5 * ntr * p.jDim * 8
            out["ring_syn_fft"] = xs + ys
            out["syn_cheb"] = ys + grid.nslots * phys1
    return out


def stage_flops(grid):
    """Useful FP64 MFMA flops per launch of the Chebyshev contractions of the FP64 form
    (RLZ / SLZ): the forward point -> mode contraction (2 b_kDim flops per input value)
    and the inverse mode -> point epilogue of the synthesis (2 b_kDim flops per output
    value of each of the 7 slots).  The ring transforms run as chirp-z FFTs on the
    vector units and carry no MFMA rate."""
    p = grid.params
    nv = grid.resolved.nvars
    geom = {"SL": "RL", "SLZ": "RLZ"}.get(grid.geometry, grid.geometry)
    if geom != "RLZ":
        return {}
    bk = p.b_kDim
    # npts already counts the kDim points of every column
    return {"syn_cheb": 2.0 * bk * 7 * grid.npts * nv, "cheb_fwd": 2.0 * bk * grid.npts * nv}


KERNEL_OF_STAGE = {"ring_syn_cheb": "k_ring_syn", "ring_syn": "k_syn_fft", "ring_syn_fft": "k_syn_fft",
                   "syn_cheb": "k_syn_cheb", "ring_ana": "k_ana_fft",
                   "sb": "k_sb_tile", "sa_si": "k_sasi", "cheb_fwd": "k_cheb_fwd_pts", "cheb_inv": "k_cheb_inv"}


def measured_traffic(workload, stage):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy), or None."""
    import json
    path = os.path.join(REPO, "profiles", f"{workload}_fp64_traffic.json")
    if not os.path.exists(path):
        return None
    k = json.load(open(path)).get("kernels", {}).get(KERNEL_OF_STAGE.get(stage, stage))
    return None if k is None else round(k["fetch_bytes"] + k["write_bytes"])


def bench_batch(args, torch, S, W):
    """Batching form: 64 RL / RLZ grids (workloads.batch_params) in batched launches replayed
    from a hipGraph, tuned launch shapes; FP64 and FP32 forms on the same draw.  Returns a dict."""
    from springsteel_hip import capi
    params, seeds, total, scale = W.batch_params()
    npv = 0
    bases = []
    for gp, sd in zip(params, seeds):
        # rl_field on the ring points (constant in z), per-grid seed
        rl = S.SpringsteelGridParameters(**{**gp.__dict__, "geometry": "RL", "jDim": 0, "b_jDim": 0,
                                            "kDim": 0, "b_kDim": 0})
        pts = S.getGridpoints(S.SpringsteelGrid(S.compute_derived_params(rl)))
        # values representable in Float32: both forms transform the same numbers
        bases.append(W.rl_field(pts, len(gp.vars), seed=sd).astype(np.float32).astype(np.float64))
    out = {"grids": len(params), "rl_grids": sum(p.geometry == "RL" for p in params),
           "num_cells_scale": round(scale, 4), "fp64_footprint_GB": round(total / 1e9, 1)}
    sample = {}
    for form, name in ((S.SSG_FP64, "fp64"), (S.SSG_FP32_REFINE, "fp32_refine")):
        grids = [S.createGrid(gp, precision=form) for gp in params]
        phys, spec, ins = [], [], []
        pdt = torch.float32 if form == S.SSG_FP32_REFINE else torch.float64   # physical storage of the form
        for g, base in zip(grids, bases):
            nv = g.resolved.nvars
            phys.append(torch.zeros((g.nslots, nv, g.npts), dtype=pdt, device="cuda"))
            spec.append(torch.zeros((nv, g.nspec), dtype=torch.float64, device="cuda"))
            b = torch.from_numpy(np.ascontiguousarray(base.T)).cuda().to(pdt)
            kz = g.params.kDim if g.geometry == "RLZ" else 1
            ins.append(b.unsqueeze(2).expand(nv, b.shape[1], kz).reshape(nv, -1).contiguous())
        npv = sum(g.npts * g.resolved.nvars for g in grids)
        batch = capi.Batch(grids, phys, spec)
        t0 = time.perf_counter()
        batch.tune()
        torch.cuda.synchronize()
        tune_s = time.perf_counter() - t0

        def load():
            for ph, x in zip(phys, ins):
                ph[0].copy_(x)
            torch.cuda.synchronize()

        load()
        for _ in range(args.warmup):
            batch.spectral()
            batch.grid()
        torch.cuda.synchronize()
        load()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.spectral()
            batch.grid()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        load()
        batch.spectral()
        batch.grid()
        torch.cuda.synchronize()
        # per-slot differences of the forms on the smallest RLZ and RL grids
        picks = [min((i for i, g in enumerate(grids) if g.geometry == geo), key=lambda i: grids[i].npts, default=None)
                 for geo in ("RLZ", "RL")]
        for i in picks:
            if i is not None:
                sample.setdefault(i, {})[name] = phys[i][:, 0].double().cpu().numpy()
        nsl, ngl = batch.launches()
        out[name] = {"value": round(npv / (ms * 1e-3) / 1e6, 2), "unit": "Mpts/s", "ms_per_step": round(ms, 3),
                     "launches_per_round_trip": nsl + ngl, "graphs_per_round_trip": 2, "tune_s": round(tune_s, 1)}
        log(f"batch {name}: {ms:.2f} ms per round trip of {npv / 1e6:.0f} M point-variables "
            f"({nsl} + {ngl} launches), tuning {tune_s:.1f} s")
        batch.close()
        del batch, grids, phys, spec, ins
        torch.cuda.empty_cache()
    out["point_vars_per_step"] = npv
    out["fp32_speedup_vs_fp64"] = round(out["fp64"]["ms_per_step"] / out["fp32_refine"]["ms_per_step"], 3)
    # FP32 form against the FP64 form (which is pinned to the oracle), variable 1, per slot
    from springsteel_hip.metrics import rel_l2_slots
    errs = {}
    for i, d in sample.items():
        a, r = d["fp32_refine"], d["fp64"]
        gp = S.compute_derived_params(params[i])
        # slot scale floors: the fields are constant in z, so the z-derivative slots hold rounding only
        errs[gp.geometry] = [float(f"{e:.3e}") for e in rel_l2_slots(a.T, r.T, gp.geometry, gp)]
    out["fp32_vs_fp64_relL2_var1"] = errs
    return out


def bench_regular(args, torch, S, grid, spec, params):
    """regularGridTransform (device form) of the grid's spectral coefficients at the
    reference's default regular points (i_regular_out x j_regular_out x k_regular_out):
    whole-call time (host tables, uploads, kernels) and per-stage device times; CPU
    baseline on a bounded sample of radii.  Returns a dict."""
    nv = grid.resolved.nvars
    p = grid.params
    axes = [a for a in S.grid._regular_axes(grid) if a is not None]
    npr = int(np.prod([len(a) for a in axes]))
    log(f"regular transform: {' x '.join(str(len(a)) for a in axes)} points, {nv} vars, {grid.nslots} slots")
    out = None
    for _ in range(max(args.warmup, 1)):
        out = S.regularGridTransform_device(grid, spec, *axes)
    torch.cuda.synchronize()
    del out
    grid.plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = S.regularGridTransform_device(grid, spec, *axes)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = grid.plan.stage_times()
    grid.plan.set_timing(False)
    res = {"value": round(npr * nv / dt / 1e6, 2), "unit": "Mpts/s (regular point-variables, all slots)",
           "ms_per_call": round(dt * 1e3, 3), "points": [len(a) for a in axes], "nvars": nv,
           "out_GB": round(npr * nv * grid.nslots * 8 / 1e9, 3),
           "stages": {n: {"ms": round(t, 4)} for n, t in st}}
    # FP64 MFMA rate of the Fourier contraction (RL / RLZ): 2 M N K per family and batch
    fr = [t for n, t in st if n == "reg_fourier"]
    if fr and grid.geometry != "RZ":
        nzb = p.b_kDim if grid.geometry in ("RLZ", "SLZ") else 1
        nks = grid.nspec // (p.b_iDim * nzb)                 # 2K + 1 spectral lines per B-spline row
        flops = 2.0 * len(axes[0]) * len(axes[1]) * nks * nv * nzb * 5
        res["fourier_TFLOPs"] = round(flops / (fr[0] * 1e-3) / 1e12, 2)
        res["fourier_gemm"] = {"M": len(axes[0]) * nv * nzb, "N": len(axes[1]), "K": nks, "families": 5}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        threads = min(len(os.sched_getaffinity(0)), 16)
        one = S.SpringsteelGridParameters(**{**params.__dict__, "vars": {"v0": 1}})
        og = O.OracleGrid(S.resolve(S.compute_derived_params(one)))
        hs = np.asfortranarray(spec[0].cpu().numpy().reshape(-1, 1))
        sub = [axes[0][:: max(1, len(axes[0]) // 8)]] + list(axes[1:])
        nsub = int(np.prod([len(a) for a in sub]))
        log(f"regular cpu baseline: oracle, variable 1, {len(sub[0])} of {len(axes[0])} radii, {threads} threads")
        t0 = time.perf_counter()
        og.regular_transform(hs, *sub, threads=threads)
        cdt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(nsub / cdt / 1e6, 3), "unit": "Mpts/s", "cores": threads, "kind": "port",
                               "sample": f"variable 1 of {nv} at {len(sub[0])} of the {len(axes[0])} radii "
                                         f"({nsub} points), C restatement with OpenMP"}
    return res


def bench_filter(args, torch, S, grid, spec, params):
    """applyFilter! (device form) on the grid's spectral array with a Fourier low pass
    (Hann taper over wavenumbers 401..500, zero above) and a Chebyshev super-Gaussian
    (sigma 12, order 4) on every variable: the kernel's average time from the plan's
    stage events and its algorithmic bytes (read + write of every scaled line, write of
    every zeroed line) against HBM peak.  The grid's filter is cleared afterwards."""
    from springsteel_hip.filters import GaussianFilter, SpectralFilter, filter_weights
    p = params
    fp = S.SpringsteelGridParameters(**{**p.__dict__,
                                        "fourier_filter": {"default": SpectralFilter(low_pass=400, window="hann",
                                                                                     taper_width=100)},
                                        "chebyshev_filter": {"default": GaussianFilter(sigma=12.0, order=4)}})
    fp = S.compute_derived_params(fp)
    names = sorted(fp.vars, key=lambda k: fp.vars[k])
    K = fp.iDim + fp.patchOffsetL
    fw, cw = filter_weights(fp, names, K, fp.b_kDim)
    grid.plan.set_filter(fw, cw)
    # the lines the library will touch, as it decides them (isapprox 1 skipped, 0 stored)
    nl = 1 if grid.geometry == "RZ" else 1 + 2 * K
    nzb = fp.b_kDim if grid.geometry in ("RLZ", "SLZ") else 1
    kk = (np.arange(nl) + 1) // 2
    w = (cw[:, :nzb, None] * fw[:, None, kk]).reshape(len(names), -1)
    rtol = np.sqrt(np.finfo(np.float64).eps)
    skip = (w == 1.0) | (np.isfinite(w) & (np.abs(w - 1.0) <= rtol * np.maximum(np.abs(w), 1.0)))
    nzero = int(np.sum(~skip & (w == 0.0)))
    nmul = int(np.sum(~skip & (w != 0.0)))
    b = fp.b_iDim
    nbytes = 8.0 * b * (2 * nmul + nzero)
    work = spec.clone()
    for _ in range(max(args.warmup, 1)):
        S.applyFilter_device(grid, work)
    torch.cuda.synchronize()
    grid.plan.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.applyFilter_device(grid, work)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = dict(grid.plan.stage_times())
    grid.plan.set_timing(False)
    grid.plan.set_filter(None, None)
    kms = st.get("filter", float("nan"))
    gbs = nbytes / (kms * 1e-3) / 1e9
    log(f"filter: {nmul} scaled + {nzero} zeroed lines of {w.size}, kernel {kms:.4f} ms, {gbs:.0f} GB/s")
    return {"ms_per_call": round(dt * 1e3, 4), "kernel_ms": round(kms, 4),
            "lines": {"scaled": nmul, "zeroed": nzero, "total": int(w.size)},
            "algorithmic_GB": round(nbytes / 1e9, 4), "achieved_GBs": round(gbs, 1),
            "hbm_frac": round(gbs / 8000.0, 3),
            "filters": "fourier: SpectralFilter(low_pass=400, window=:hann, taper_width=100); "
                       "chebyshev: GaussianFilter(sigma=12, order=4); all variables"}


def bench_unstructured(args, torch, S, grid, spec, params, npts=1_000_000):
    """evaluate_unstructured (device form) of the grid's spectral coefficients at npts
    seeded points uniform over the domain: whole-call time (host sort and tables,
    uploads, kernels) and per-stage device times; CPU baseline on a bounded sample of
    the points.  Returns a dict."""
    nv = grid.resolved.nvars
    p = grid.params
    rng = np.random.default_rng(11)
    cols = [rng.uniform(p.iMin, p.iMax, npts)]
    if grid.geometry != "RZ":
        cols.append(rng.uniform(0.0, 2 * np.pi, npts))
    if grid.geometry not in ("RL", "SL"):
        cols.append(rng.uniform(p.kMin, p.kMax, npts))
    pts = np.column_stack(cols)
    log(f"unstructured evaluation: {npts} points, {nv} vars")
    out = None
    for _ in range(max(args.warmup, 1)):
        out = S.evaluate_unstructured_device(grid, spec, pts)
    torch.cuda.synchronize()
    del out
    grid.plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = S.evaluate_unstructured_device(grid, spec, pts)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = grid.plan.stage_times()
    grid.plan.set_timing(False)
    res = {"value": round(npts * nv / dt / 1e6, 3), "unit": "Mpts/s (point-variables)", "ms_per_call": round(dt * 1e3, 3),
           "points": npts, "nvars": nv, "stages": {n: {"ms": round(t, 4)} for n, t in st}}
    ev = [t for n, t in st if n == "unstr_eval"]
    if ev and grid.geometry != "RZ":
        nzb = p.b_kDim if grid.geometry in ("RLZ", "SLZ") else 1
        nks = grid.nspec // (p.b_iDim * nzb)
        flops = 2.0 * npts * 4 * nks * nv * nzb            # the (tap, line) products of every point
        res["eval_TFLOPs"] = round(flops / (ev[0] * 1e-3) / 1e12, 2)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        threads = min(len(os.sched_getaffinity(0)), 16)
        one = S.SpringsteelGridParameters(**{**params.__dict__, "vars": {"v0": 1}})
        og = O.OracleGrid(S.resolve(S.compute_derived_params(one)))
        hs = np.asfortranarray(spec[0].cpu().numpy().reshape(-1, 1))
        nsub = 20000
        log(f"unstructured cpu baseline: oracle, variable 1, {nsub} of the points, {threads} threads")
        t0 = time.perf_counter()
        og.evaluate_unstructured(hs, pts[:nsub], threads=threads)
        cdt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(nsub / cdt / 1e6, 4), "unit": "Mpts/s", "cores": threads, "kind": "port",
                               "sample": f"variable 1 of {nv} at {nsub} of the {npts} points, C restatement with OpenMP"}
    return res


def cpu_worker(args):
    """One oracle round trip of variable 1 of the workload on args.cpu_worker threads;
    prints {"seconds", "points"}.  Imports no GPU runtime."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import springsteel_hip.params as P
    from springsteel_hip import workloads as W
    from springsteel_hip.grid import SpringsteelGrid
    params = W.WORKLOADS[args.workload]["params"]()
    one = P.SpringsteelGridParameters(**{**params.__dict__, "vars": {"v0": 1}})
    g = SpringsteelGrid(P.compute_derived_params(one))
    field = W.workload_input(args.workload, g)
    og = O.OracleGrid(g.resolved)
    hp = og.new_physical()
    hp[:, 0, 0] = field[:, 0]
    t0 = time.perf_counter()
    spec = og.spectral_transform(hp, threads=args.cpu_worker)
    og.grid_transform(spec, threads=args.cpu_worker)
    dt = time.perf_counter() - t0
    print(json.dumps({"seconds": dt, "points": g.npts, "threads": args.cpu_worker}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rlz_vortex")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--parity", action="store_true", help="full-size parity of variable 1 against the oracle")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 form")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in launch shapes")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-grid batching form")
    ap.add_argument("--no-regular", action="store_true", help="skip the regular-grid evaluation leg")
    ap.add_argument("--no-unstructured", action="store_true", help="skip the unstructured evaluation leg")
    ap.add_argument("--no-filter", action="store_true", help="skip the applyFilter! leg")
    ap.add_argument("--cpu-worker", type=int, default=0, metavar="THREADS",
                    help="internal: time one oracle round trip of variable 1 on THREADS host threads and print "
                         "one JSON line (touches no GPU)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the transform path "
              "here is single-GPU; multi-GPU sharding is a design, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    if args.cpu_worker:
        cpu_worker(args)
        return
    # the single-threaded CPU baseline (the reference's own threading) runs in a child
    # process beside the GPU legs; it is started before this process touches the GPU
    worker = None
    if not args.no_cpu and args.workload != "batch64":
        import subprocess
        env = {**os.environ, "OMP_NUM_THREADS": "1"}
        worker = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", "1",
                                   "--workload", args.workload], stdout=subprocess.PIPE, env=env, text=True)

    import torch
    import springsteel_hip as S
    from springsteel_hip import workloads as W

    torch.cuda.set_device(0)
    if args.workload == "batch64":
        res = bench_batch(args, torch, S, W)
        line = {"metric": "batched RL/RLZ round-trip Mpts/s (64 grids, one hipGraph per direction), FP32 form",
                "value": res["fp32_refine"]["value"], "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": res["fp32_refine"]["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded fields)",
                "config": {"workload": "batch64"}, "batch": res}
        print(json.dumps(line), flush=True)
        return
    params = W.WORKLOADS[args.workload]["params"]()
    log(f"workload {args.workload}: creating plan")
    grid = S.createGrid(params)
    p = grid.params
    nv = grid.resolved.nvars
    log(f"npts/var {grid.npts}, nvars {nv}, nspec/var {grid.nspec}; building input")
    # values representable in Float32, so that the FP64 and FP32 forms (the latter stores
    # physical in Float32) transform the same numbers
    field = np.asfortranarray(W.workload_input(args.workload, grid).astype(np.float32).astype(np.float64))
    phys = torch.zeros((grid.nslots, nv, grid.npts), dtype=torch.float64, device="cuda")
    spec = torch.zeros((nv, grid.nspec), dtype=torch.float64, device="cuda")
    phys[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
    torch.cuda.synchronize()

    def step(g=grid, ph=phys):
        S.spectralTransform_device(g, ph, spec)
        S.gridTransform_device(g, spec, ph)

    def time_form(g, ph):
        """tune the launch shapes on these arrays, reload the input, warmup, then K timed
        steps between synchronisations; (ms per step, stage times)"""
        if not args.no_tune:
            from springsteel_hip import capi
            capi.plan_tune(g.plan, ph, spec)
            ph[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
            torch.cuda.synchronize()
        for _ in range(args.warmup):
            step(g, ph)
        torch.cuda.synchronize()
        g.plan.set_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(g, ph)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = g.plan.stage_times()
        g.plan.set_timing(False)
        return dt / args.steps * 1e3, st

    if args.copy_calib:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log(f"FP64 form: warmup {args.warmup}, timing {args.steps} steps")
    ms, stages = time_form(grid, phys)
    pts_per_step = grid.npts * nv
    value = pts_per_step / (ms * 1e-3) / 1e6
    g32 = phys32 = None
    if grid.geometry != "RZ" and not args.no_fp32:
        # FP32 form: its own plan and Float32 physical array, the same spectral array
        phys32 = torch.zeros((grid.nslots, nv, grid.npts), dtype=torch.float32, device="cuda")
        phys32[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
        g32 = S.createGrid(params, precision=S.SSG_FP32_REFINE)
        log(f"FP32 form: warmup {args.warmup}, timing {args.steps} steps")
        ms32, stages32 = time_form(g32, phys32)
    sb = stage_bytes(grid)
    sf = stage_flops(grid)
    dom_name, dom_ms = max(stages, key=lambda x: x[1])
    stage_report = {}
    for n, t in stages:
        e = {"ms": round(t, 4), "GBps": round(sb.get(n, 0) / (t * 1e-3) / 1e9, 1)}
        if n in sf:
            e["TFLOPs"] = round(sf[n] / (t * 1e-3) / 1e12, 2)
        stage_report[n] = e
    # The transform path is HBM-bound by design (every stage streams its arrays once), so
    # the roofline of the dominant kernel is its algorithmic bytes per launch (the stage's
    # compulsory input + output arrays, stage_bytes) over its average launch time against
    # HBM peak.  Its FP64 MFMA rate is reported beside it.
    achieved = sb[dom_name] / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": KERNEL_OF_STAGE.get(dom_name, dom_name), "stage": dom_name,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes": int(sb[dom_name]),
            "launch_ms": round(dom_ms, 4)}
    roof["traffic"] = measured_traffic(args.workload, dom_name)
    if dom_name in sf:
        roof["mfma_fp64_TFLOPs"] = round(sf[dom_name] / (dom_ms * 1e-3) / 1e12, 2)
        roof["mfma_fp64_frac"] = round(roof["mfma_fp64_TFLOPs"] / FP64_PEAK_TFLOPS, 4)
    # whole round trip: compulsory bytes of the boundary arrays only (read slot 1, write
    # every slot, write + read spectral), as the headline's % of HBM peak
    rt_bytes = 8.0 * nv * (grid.npts * (1 + grid.nslots) + 2 * grid.nspec)
    alg_bytes = sum(sb[n] for n, _ in stages if n in sb)
    log(f"{ms:.3f} ms/step; stages {stage_report}")

    result = {
        "metric": "RLZ round-trip Mpts/s (point-variables per second), FP64 plain form",
        "value": round(value, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded fields)",
        "config": {"workload": args.workload, "geometry": grid.geometry, "num_cells": p.num_cells,
                   "iDim": p.iDim, "kDim": p.kDim, "b_kDim": p.b_kDim, "nvars": nv,
                   "points_per_var": grid.npts, "point_vars_per_step": pts_per_step, "form": "fp64"},
        "roofline": roof,
        "stages": stage_report,
        "round_trip_hbm": {"alg_bytes": int(rt_bytes), "achieved_GBps": round(rt_bytes / (ms * 1e-3) / 1e9, 1),
                           "frac": round(rt_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes": "8 B x nvars x (npts x (1 + nslots) + 2 nspec)"},
        "stage_bytes_sum_GBps": round(alg_bytes / (ms * 1e-3) / 1e9, 1),
    }
    if not args.no_tune:
        from springsteel_hip import capi
        result["tuned_shapes"] = capi.plan_tuning(grid.plan)
    if g32 is not None:
        result["fp32_refine"] = {
            "value": round(pts_per_step / (ms32 * 1e-3) / 1e6, 2), "unit": "Mpts/s", "ms_per_step": round(ms32, 4),
            "speedup_vs_fp64": round(ms / ms32, 3),
            "dtype": "f32 storage of physical; FP64 forward path, solve and evaluation; FP32 synthesis",
            "stages": {n: {"ms": round(t, 4)} for n, t in stages32}}
        if not args.no_tune:
            from springsteel_hip import capi
            result["fp32_refine"]["tuned_shapes"] = capi.plan_tuning(g32.plan)
        # accuracy of the FP32 form against the FP64 form on this grid (the FP64 form is
        # held to the oracle at 1e-12 by the tests), per variable and derivative order
        from springsteel_hip.metrics import rel_l2_slots_device
        inp = torch.from_numpy(np.ascontiguousarray(field.T)).cuda()
        phys[0].copy_(inp)
        step(grid, phys)
        torch.cuda.synchronize()
        phys32[0].copy_(inp)
        step(g32, phys32)
        torch.cuda.synchronize()
        orders = {"d0": (0,), "d1": (1, 3, 5), "d2": (2, 4, 6)} if grid.nslots == 7 else \
            {"d0": (0,), "d1": (1, 3), "d2": (2, 4)}
        per_var = []
        for v in range(nv):
            e = rel_l2_slots_device(phys32[:, v].double(), phys[:, v], grid.geometry, grid.params)
            per_var.append({o: float(f"{max(e[s] for s in ss):.3e}") for o, ss in orders.items()})
        worst = {o: max(pv[o] for pv in per_var) for o in orders}
        result["fp32_refine"]["vs_fp64_relL2"] = {"per_variable": per_var, "worst": worst,
                                                  "meets_1e-6": bool(max(worst.values()) <= 1e-6)}
        log(f"FP32 form vs FP64 form: worst rel L2 {worst}")
        del inp

    if worker is not None:
        log("waiting for the single-threaded CPU baseline")
        out, _ = worker.communicate()
        try:
            w = json.loads(out.strip().splitlines()[-1])
            result["cpu_baseline"] = {
                "value": round(w["points"] / w["seconds"] / 1e6, 4), "unit": "Mpts/s", "cores": 1, "kind": "port",
                "seconds": round(w["seconds"], 2),
                "sample": f"one round trip (spectralTransform + gridTransform) of variable 1 of {nv} "
                          f"({w['points']} points) of the same grid on one host thread, the reference's own "
                          f"threading; C restatement of the reference's loops, run beside the GPU legs"}
        except Exception as e:   # report, keep the main line
            result["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu or args.parity:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        threads = min(len(os.sched_getaffinity(0)), 16)
        one = S.SpringsteelGridParameters(**{**params.__dict__, "vars": {"v0": 1}})
        og = O.OracleGrid(S.resolve(S.compute_derived_params(one)))
        hp = og.new_physical()
        hp[:, 0, 0] = field[:, 0]
        log(f"cpu baseline: oracle, variable 1 of {nv}, {threads} threads")
        t0 = time.perf_counter()
        ospec = og.spectral_transform(hp, threads=threads)
        ophys = og.grid_transform(ospec, threads=threads)
        cdt = time.perf_counter() - t0
        if not args.no_cpu:
            result["cpu_baseline_threads"] = {
                "value": round(grid.npts / cdt / 1e6, 3), "unit": "Mpts/s", "cores": threads, "kind": "port",
                "sample": f"one round trip of variable 1 of {nv} ({grid.npts} points) of the same grid, "
                          f"C restatement with OpenMP over lines and columns (not the reference's threading)"}
        if args.parity:
            # slot 1 was overwritten by the timed round trips: restore the input first
            phys[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
            step()
            torch.cuda.synchronize()
            from springsteel_hip.metrics import rel_l2_slots
            gspec = spec[0].cpu().numpy()
            errs = {"spectral": float(np.linalg.norm(gspec - ospec[:, 0]) / np.linalg.norm(ospec[:, 0]))}
            gphys = np.stack([phys[s, 0].cpu().numpy() for s in range(grid.nslots)], axis=1)
            for s, e in enumerate(rel_l2_slots(gphys, ophys[:, 0, :], grid.geometry, grid.params)):
                errs[f"slot{s + 1}"] = e
            result["parity_var1_relL2"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
            if g32 is not None:
                phys32[0].copy_(torch.from_numpy(np.ascontiguousarray(field.T)))
                step(g32, phys32)
                torch.cuda.synchronize()
                gspec = spec[0].cpu().numpy()
                e32 = {"spectral": float(np.linalg.norm(gspec - ospec[:, 0]) / np.linalg.norm(ospec[:, 0]))}
                gphys = np.stack([phys32[s, 0].double().cpu().numpy() for s in range(grid.nslots)], axis=1)
                for s, e in enumerate(rel_l2_slots(gphys, ophys[:, 0, :], grid.geometry, grid.params)):
                    e32[f"slot{s + 1}"] = e
                result["fp32_refine"]["parity_var1_relL2"] = {k: float(f"{v:.3e}") for k, v in e32.items()}
    if not args.no_regular:
        try:
            result["regular"] = bench_regular(args, torch, S, grid, spec, params)
        except Exception as e:   # report, keep the main line
            result["regular"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_unstructured:
        try:
            result["unstructured"] = bench_unstructured(args, torch, S, grid, spec, params)
        except Exception as e:   # report, keep the main line
            result["unstructured"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_filter and grid.geometry in ("RLZ", "SLZ"):
        try:
            result["filter"] = bench_filter(args, torch, S, grid, spec, params)
        except Exception as e:   # report, keep the main line
            result["filter"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_batch and args.workload == "rlz_vortex":
        # the batching form on 64 grids, after the single-grid arrays are released
        del phys, spec, phys32
        grid.close()
        if g32 is not None:
            g32.close()
        torch.cuda.empty_cache()
        try:
            result["batch64"] = bench_batch(args, torch, S, W)
        except Exception as e:   # report, keep the main line
            result["batch64"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
