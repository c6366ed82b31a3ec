"""Write the golden fixtures of the transform path: small RL, RZ and RLZ grids,
seeded inputs, and the oracle's spectral and physical outputs.

The oracle (oracle/ssg_oracle.c) is pinned independently by
tests/test_oracle_pins.py; these fixtures freeze its outputs so that the HIP
path and any later change of the oracle are checked against the same numbers.

  python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "springsteel.jl_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))

# name -> SpringsteelGridParameters keyword arguments (BCs as legacy dict names)
CASES = {
    "rl_nc10": dict(geometry="RL", num_cells=10, iMin=0.0, iMax=100.0, vars={"u": 1}),
    "rl_nc8_two_vars_bcs": dict(geometry="RL", num_cells=8, iMin=0.0, iMax=50.0, vars={"u": 1, "v": 2},
                                BCL={"u": "R1T0", "v": "R0"}, BCR={"u": "R0", "v": "R1T1"},
                                max_wavenumber={"v": 5}),
    "rz_nc8_kdim15": dict(geometry="RZ", num_cells=8, iMin=0.0, iMax=40.0, kMin=0.0, kMax=20.0, kDim=15,
                          vars={"u": 1, "w": 2}, BCB={"default": "R0", "w": "R1T0"},
                          BCT={"default": "R0", "w": "R1T1"}),
    "rlz_nc4_kdim8": dict(geometry="RLZ", num_cells=4, iMin=0.0, iMax=60.0, kMin=0.0, kMax=10.0, kDim=8,
                          vars={"u": 1, "v": 2}),
}


def params_of(kw):
    import springsteel_hip as S
    kw = dict(kw)
    for key, mod in (("BCL", S.CubicBSpline), ("BCR", S.CubicBSpline), ("BCB", S.Chebyshev), ("BCT", S.Chebyshev)):
        if key in kw:
            kw[key] = {v: getattr(mod, name) for v, name in kw[key].items()}
    return S.SpringsteelGridParameters(**kw)


def main():
    import oracle as O
    import springsteel_hip as S
    from parity_util import seeded_field
    from springsteel_hip.params import compute_derived_params, resolve
    for seed, (name, kw) in enumerate(CASES.items()):
        gp = compute_derived_params(params_of(kw))
        og = O.OracleGrid(resolve(gp))
        pts = og.gridpoints()
        nv = len(gp.vars)
        phys = og.new_physical()
        phys[:, :, 0] = seeded_field(pts, nv, 100 + seed, gp.geometry)
        spec = og.spectral_transform(phys, threads=1)
        out = og.grid_transform(spec, threads=1)
        np.savez(os.path.join(HERE, f"{name}.npz"), params=np.array(json.dumps(kw)), input=phys[:, :, 0],
                 spectral=spec, physical=out)
        print(name, phys.shape, spec.shape)


if __name__ == "__main__":
    main()
