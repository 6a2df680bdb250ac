"""Edge cases of the step on the GPU against the oracle.

* Dense cells (h = 2 dr: ~64 particles per cell, ~1700 candidates and ~270
  neighbours per particle): every chunked candidate loop runs past its first
  32-candidate chunk and the pair lists overflow into their cell-scan
  fallback; all interaction variants stay bit-identical and one step
  matches the oracle.
* Walls only (no soil): the moving wall advances exactly as in the oracle.
* One soil particle and no walls: free fall, bit-identical to the FP32
  oracle.
"""
import sys

import numpy as np
import pytest

from conftest import relerr
from retainwall_code_amd import _abi, cases, sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a


def subset(case, keep_bnd, keep_soil, name):
    """The case restricted to the given wall and soil rows (format keeps walls first)."""
    rows = np.concatenate([np.asarray(keep_bnd, int), case.npb + np.asarray(keep_soil, int)])
    p = dict(case.params)
    p["np"], p["npb"], p["npm"] = len(rows), len(keep_bnd), len(keep_soil)
    return cases.Case(p, case.codes[rows], case.pospres[rows], case.velrhop[rows], case.idmassidx[rows],
                      case.str[rows], case.stn[rows], case.normal[np.asarray(keep_bnd, int)], case.motion,
                      case.probe_box, name)


def oracle_from(ctx, O, dtype, case):
    o = O.Oracle(ctx.get_params(), dtype, ctx.get_motion(len(case.motion)))
    names = ["codes", "cellpos", "orig_id", "pos", "pos_ori", "vel", "idm", "normal", "str", "stn", "cvar",
             "cmag", "pos_ini", "vel_ini", "str_ini", "stn_ini", "accad", "spin", "acestn", "cell_beg", "cell_end"]
    o.load_fields(ctx.snapshot(names))
    o.time, o.step = ctx.time()
    return o


def dense_case():
    return cases.retaining_wall(box_length=0.05, box_width=0.03, box_height=0.04, dr=0.0025, hdr=2.0,
                                fix_wall_height=0.01, wall_extra=0.01, motion=("rotate", -0.01), jitter=0.05,
                                void_range=(0.65, 0.75), seed=21, name="dense")


def test_dense_cells_variants_identical():
    case = dense_case()
    say(f"[dense] np={case.np} npb={case.npb} npm={case.npm}")
    outs = []
    for v in [_abi.INTERACTION_NEIGHBOR_LIST, _abi.INTERACTION_REFERENCE_ORDER, _abi.INTERACTION_TILE,
              _abi.INTERACTION_PAIR_LIST]:
        ctx = sim.open_case(case, variant=v)
        ctx.step(2)
        ctx.sync()
        if v == _abi.INTERACTION_PAIR_LIST:
            ov, _, _ = ctx.pair_list_status()
            assert ov == 1  # ~270 neighbours per particle exceed the 160-entry lists
        outs.append(ctx.snapshot(["accad", "acestn", "spin", "vel", "str", "cmag", "cvar", "pos"]))
        ctx.close()
        say(f"  variant {v} done")
    for o in outs[1:]:
        for k in outs[0]:
            np.testing.assert_array_equal(bits(outs[0][k]), bits(o[k]), err_msg=k)


def test_dense_cells_step_matches_oracle(oracle_mod):
    case = dense_case()
    ctx = sim.open_case(case)
    ctx.step(1)
    ctx.sync()
    o32 = oracle_from(ctx, oracle_mod, np.float32, case)
    o32.run(1)
    ctx.step(1)
    ctx.sync()
    soil = slice(case.npb, case.np)
    for k in ["pos", "vel", "str", "idm"]:
        e = relerr(ctx.get(k)[soil], o32[k][soil])
        say(f"  dense {k} relerr vs FP32 oracle {e:.2e}")
        assert e <= 2e-6, k
    ctx.close()


def test_walls_only(oracle_mod):
    base = cases.config("tiny")
    case = subset(base, np.arange(base.npb), [], "walls")
    say(f"[walls only] np={case.np}")
    ctx = sim.open_case(case)
    assert ctx.counts() == (case.npb, 0, 0)
    o32 = oracle_from(ctx, oracle_mod, np.float32, case)
    o32.run(3)
    ctx.step(3)
    ctx.sync()
    g = ctx.snapshot(["pos", "cellpos", "normal", "vel", "orig_id"])
    np.testing.assert_array_equal(g["cellpos"], o32["cellpos"])
    np.testing.assert_array_equal(g["orig_id"], o32["orig_id"])
    for k in ["pos", "normal", "vel"]:
        e = relerr(g[k], o32[k])
        say(f"  walls {k} relerr vs FP32 oracle {e:.2e}")
        assert e <= 2e-6, k
    t, s = ctx.time()
    assert s == o32.step and abs(t - o32.time) < 1e-15
    ctx.close()


def test_single_soil_particle_free_fall(oracle_mod):
    base = cases.config("tiny")
    case = subset(base, [], [base.npm // 2], "one")
    say("[one particle] free fall, 3 steps")
    ctx = sim.open_case(case)
    o32 = oracle_from(ctx, oracle_mod, np.float32, case)
    o32.run(3)
    ctx.step(3)
    ctx.sync()
    for k in ["pos", "vel", "str", "idm"]:
        e = relerr(ctx.get(k), o32[k])
        say(f"  one particle {k} relerr vs FP32 oracle {e:.2e}")
        assert e <= 2e-6, k
    # no neighbours: constant acceleration g, so v = g t whatever the integrator
    t, _ = ctx.time()
    gz = case.params["gravity"][2]
    v = ctx.get("vel")[0]
    assert v[0] == 0.0 and v[1] == 0.0
    assert abs(v[2] - gz * t) <= 1e-5 * abs(gz * t), (v[2], gz * t)
    ctx.close()
