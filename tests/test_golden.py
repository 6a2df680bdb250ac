"""The committed c1 wall-force histories (tests/golden/c1_wall_force.json,
written by tests/golden/make_c1_wall_force.py) still describe the generated
c1 cases, and the FP64 oracle reproduces their first sample bit for bit.

Workload: 0.4 x 0.2 x 0.3 m backfill at dx 5 mm behind a wall translating
away at 1 mm/s (the wall force of utilities.cu:578-670 every 50 steps).
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_c1_wall_force as G  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "c1_wall_force.json")


@pytest.fixture(scope="module")
def golden():
    with open(FIXTURE) as fh:
        return json.load(fh)


def test_fixture_shape(golden):
    assert golden["steps"] == G.STEPS and golden["every"] == G.EVERY
    for lat in G.LATTICES:
        row = golden["lattices"][lat]
        for k in ("f64", "f32"):
            f = np.asarray(row[k])
            assert f.shape == (G.STEPS // G.EVERY,) and np.all(np.isfinite(f))
        # the FP32 run stays near the FP64 run (within the rounding-driven
        # contact flips measured on this workload, a few per cent)
        f64, f32 = np.asarray(row["f64"]), np.asarray(row["f32"])
        assert np.max(np.abs(f32 - f64)) / np.max(np.abs(f64)) < 0.1


@pytest.mark.parametrize("lattice", list(G.LATTICES))
def test_fixture_matches_generated_case(golden, lattice):
    case = G.case_for(lattice)
    row = golden["lattices"][lattice]
    assert (case.np, case.npb) == (row["np"], row["npb"])
    assert G.case_digest(case) == row["case_sha256"]
    assert [float(x) for x in case.probe_box] == row["probe_box"]


@pytest.mark.parametrize("lattice", list(G.LATTICES))
def test_oracle_reproduces_first_sample(golden, lattice, oracle_mod):
    _, f, _ = G.run(lattice, np.float64, G.EVERY, G.EVERY)
    assert float(f[0]) == golden["lattices"][lattice]["f64"][0]
