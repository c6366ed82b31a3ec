"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
small RL / RZ / RLZ grids with seeded inputs and the oracle's outputs.

CPU: the oracle still reproduces every fixture exactly.
GPU: the HIP path reproduces every fixture within the FP64 parity bar
(relative L2 per variable and slot, slot scale floors, sensitivity margin)."""
import glob
import json
import os

import numpy as np
import pytest

import oracle as O
import springsteel_hip as S
from springsteel_hip.params import compute_derived_params, resolve

from golden.make_golden import params_of
from parity_util import TOL_FP64, errors, sensitivity, slot_tolerances

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
IDS = [os.path.splitext(os.path.basename(f))[0] for f in FIXTURES]


def _load(path):
    z = np.load(path)   # allow_pickle stays False: plain arrays and a JSON string
    return json.loads(str(z["params"])), z["input"], z["spectral"], z["physical"]


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    kw, inp, spec, phys = _load(path)
    og = O.OracleGrid(resolve(compute_derived_params(params_of(kw))))
    hp = og.new_physical()
    hp[:, :, 0] = inp
    ospec = og.spectral_transform(hp, threads=1)
    ophys = og.grid_transform(ospec, threads=1)
    assert np.array_equal(ospec, spec)
    assert np.array_equal(ophys, phys)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=IDS)
def test_hip_reproduces_fixture(path):
    kw, inp, spec, phys = _load(path)
    grid = S.createGrid(params_of(kw))
    grid.physical[:, :, 0] = inp
    S.spectralTransform_(grid)
    S.gridTransform_(grid)
    es = errors(grid.spectral, spec)
    ep = errors(grid.physical, phys, grid.geometry, grid.params)
    og = O.OracleGrid(grid.resolved)
    hp = og.new_physical()
    hp[:, :, 0] = inp
    tols = slot_tolerances(sensitivity(og, hp, phys, grid.geometry, grid.params))
    print(f"\n{os.path.basename(path)}: spectral {max(es):.2e}, slots "
          f"{[f'{max(e[s] for e in ep):.1e}' for s in range(phys.shape[2])]}")
    assert max(es) <= TOL_FP64, es
    for v in range(len(ep)):
        for s in range(phys.shape[2]):
            assert ep[v][s] <= tols[v][s], (v, s, ep[v], tols[v])
    grid.close()
