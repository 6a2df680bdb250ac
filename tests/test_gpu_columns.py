"""The COLUMNS interaction variant (h-sized sub-cell windows).

* neighbour sets: per soil particle the same original ids as the 27-cell
  scan, on shuffled small cases and on the full-size c2 workload;
* one interaction from the same state against the FP32 and FP64 oracles
  (the summation order differs from the reference's, so against both the
  FP64 tolerances apply), small case and c2;
* whole steps against the FP64 oracle's steps, and graph replay equal to
  eager launches;
* steps with the predictor and corrector fused into the interaction
  (lq_set_fusion) bit-identical to the unfused order.
"""
import numpy as np
import pytest

from conftest import relerr
from retainwall_code_amd import _abi, cases, sim
from test_gpu_parity import TOL_F64_RATE, TOL_F64_STATE, bits, oracle_from, say, shuffled

pytestmark = pytest.mark.gpu
COLS = _abi.INTERACTION_COLUMNS


def _sets_equal(ctx_a, ctx_b, max_nb=200):
    ca, ia = ctx_a.neighbor_lists(max_nb)
    cb, ib = ctx_b.neighbor_lists(max_nb)
    np.testing.assert_array_equal(ca, cb)
    assert ca.max() <= max_nb
    ia = np.where(np.arange(max_nb)[None, :] < ca[:, None], ia, -1)
    ib = np.where(np.arange(max_nb)[None, :] < cb[:, None], ib, -1)
    np.testing.assert_array_equal(np.sort(ia, axis=1), np.sort(ib, axis=1))
    return int(ca.sum())


@pytest.mark.parametrize("name,shuffle", [("tiny", True), ("small", True), ("small", False)])
def test_columns_neighbor_sets(name, shuffle):
    case = cases.config(name)
    if shuffle:
        case = shuffled(case, 4)
    say(f"[columns sets] case={case.name}")
    a = sim.open_case(case, variant=_abi.INTERACTION_NEIGHBOR_LIST)
    b = sim.open_case(case, variant=COLS)
    for _ in range(2):
        a.step(3)
        b.set_interaction(_abi.INTERACTION_NEIGHBOR_LIST)
        b.step(3)  # same trajectory bit for bit, then compare the searches
        b.set_interaction(COLS)
        a.sync()
        b.sync()
        n = _sets_equal(a, b)
        say(f"  {n} pairs, sets identical")
    a.close()
    b.close()


def _interaction_vs_oracles(case, oracle_mod, steps):
    ctx = sim.open_case(case, variant=COLS)
    ctx.step(steps)
    ctx.step_calculation()
    ctx.neighbor_search()
    ctx.sort_arrays_primary()
    ctx.rigid_body_update()
    ctx.adami()
    ctx.find_corner()
    ctx.sync()
    o32 = oracle_from(ctx, oracle_mod, np.float32, case)
    o64 = oracle_from(ctx, oracle_mod, np.float64, case)
    o32.interaction()
    o64.interaction()
    ctx.interaction()
    ctx.sync()
    soil = slice(case.npb, case.np)
    for k in ("accad", "acestn", "spin", "cmag"):
        g = ctx.get(k)[soil]
        e32, e64 = relerr(g, o32[k][soil]), relerr(g, o64[k][soil])
        say(f"  {k:7s} relerr f32={e32:.2e} f64={e64:.2e}")
        assert e32 <= TOL_F64_RATE and e64 <= TOL_F64_RATE, k
    np.testing.assert_array_equal(ctx.get("cvar")[soil][:, 2], o32["cvar"][soil][:, 2])  # slip flags
    ctx.close()


def test_columns_interaction_small(oracle_mod):
    say("[columns interaction] case=small")
    _interaction_vs_oracles(cases.config("small"), oracle_mod, 4)


@pytest.mark.timeout(300)
def test_columns_c2_sets_and_interaction(oracle_mod):
    c2 = cases.config("c2")
    say(f"[columns c2] np={c2.np}")
    a = sim.open_case(c2, variant=_abi.INTERACTION_NEIGHBOR_LIST)
    b = sim.open_case(c2, variant=COLS)
    n = _sets_equal(a, b, 160)
    say(f"  {n} pairs, sets identical")
    a.close()
    b.close()
    _interaction_vs_oracles(c2, oracle_mod, 2)


def test_columns_whole_steps(oracle_mod):
    case = cases.config("small")
    say("[columns steps] case=small, 2 + 6 steps")
    ctx = sim.open_case(case, variant=COLS)
    ctx.step(2)
    ctx.sync()
    o64 = oracle_from(ctx, oracle_mod, np.float64, case)
    o64.run(6)
    ctx.step(6)
    ctx.sync()
    soil = slice(case.npb, case.np)
    np.testing.assert_array_equal(ctx.get("orig_id"), o64["orig_id"])
    for k in ("pos", "vel", "str"):
        e = relerr(ctx.get(k)[soil], o64[k][soil])
        say(f"  {k} relerr vs FP64 oracle {e:.2e}")
        assert e <= TOL_F64_STATE, k
    ctx.close()


def test_columns_graph_equals_eager():
    case = cases.config("small")
    a = sim.open_case(case, variant=COLS)
    b = sim.open_case(case, variant=COLS)
    a.step(4)
    b.step(4, graph=True)
    a.sync()
    b.sync()
    for k in ("pos", "vel", "str", "cvar", "cmag"):
        np.testing.assert_array_equal(bits(a.get(k)), bits(b.get(k)), err_msg=k)
    a.close()
    b.close()


FUSED_FIELDS = ["pos", "vel", "str", "stn", "idm", "cellpos", "orig_id", "cvar", "cmag", "pos_ori", "accad",
                "acestn", "spin", "pos_ini", "vel_ini", "str_ini", "stn_ini"]


@pytest.mark.parametrize("name,graph", [("small", False), ("small", True), ("tiny", True)])
def test_fused_steps_bit_identical(name, graph):
    """COLUMNS steps with the predictor / corrector fused into the interaction
    equal the unfused order bit for bit (deferred contact particles finish
    after pass B, which reads the first-layer positions saved before the
    update)."""
    case = cases.config(name)
    say(f"[fused] case={name} graph={graph}")
    a = sim.open_case(case, variant=COLS)
    b = sim.open_case(case, variant=COLS)
    a.set_fusion(True)
    b.set_fusion(False)
    for _ in range(3):
        a.step(4, graph=graph)
        b.step(4, graph=graph)
        a.sync()
        b.sync()
        for k in FUSED_FIELDS:
            np.testing.assert_array_equal(bits(a.get(k)), bits(b.get(k)), err_msg=k)
        assert a.counts() == b.counts()
    d = int(np.count_nonzero(a.get("cvar")[case.npb:, 1] > float(a.get_params().dr)))
    say(f"  12 steps bit-identical ({d} soil particles beyond dr of a wall)")
    a.close()
    b.close()


@pytest.mark.timeout(300)
def test_fused_c2_bit_identical():
    c2 = cases.config("c2")
    a = sim.open_case(c2, variant=COLS)
    b = sim.open_case(c2, variant=COLS)
    a.set_fusion(True)
    b.set_fusion(False)
    a.step(6, graph=True)
    b.step(6, graph=True)
    a.sync()
    b.sync()
    for k in FUSED_FIELDS:
        np.testing.assert_array_equal(bits(a.get(k)), bits(b.get(k)), err_msg=k)
    say("  c2: 6 fused graph steps bit-identical to unfused")
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["small", "c2"])
def test_columns_builds_bit_identical(name):
    """The two COLUMNS register builds evaluate the same pairs in the same
    order: one interaction from the same state and whole graph steps are
    bit-identical."""
    case = cases.config(name)
    say(f"[columns builds] case={name}")
    ctx = sim.open_case(case, variant=COLS)
    ctx.step(2)
    ctx.step_calculation()
    ctx.neighbor_search()
    ctx.sort_arrays_primary()
    ctx.rigid_body_update()
    ctx.adami()
    ctx.find_corner()
    ctx.sync()
    saved = {k: ctx.get(k) for k in ("cvar", "cmag")}
    outs = []
    for build in range(2):
        for k, v in saved.items():
            ctx.set(k, v)
        ctx.set_columns_build(build)
        ctx.interaction()
        ctx.sync()
        outs.append({k: ctx.get(k) for k in ("accad", "acestn", "spin", "cvar", "cmag")})
    for build in range(1, 2):
        for k in outs[0]:
            np.testing.assert_array_equal(bits(outs[build][k]), bits(outs[0][k]), err_msg=f"build {build} {k}")
    ctx.close()
    runs = []
    for build in (0, 1):
        c = sim.open_case(case, variant=COLS)
        c.set_columns_build(build)
        c.step(4, graph=True)
        c.sync()
        runs.append({k: c.get(k) for k in ("pos", "vel", "str", "cvar")})
        c.close()
    for k in runs[0]:
        np.testing.assert_array_equal(bits(runs[1][k]), bits(runs[0][k]), err_msg=k)
    say("  builds 0 and 1 bit-identical, one call and graph steps")



@pytest.mark.parametrize("name", ["small", "c2", "shepard"])
def test_columns_corner_beside_bit_identical(name):
    """COLUMNS steps with Adami and the corner search on a second stream
    beside the column build and the interaction (wall records written after
    Adami, the wall contact of near-wall particles finished after the
    search) equal the serial order bit for bit, in graph replay and eager
    steps; also with the Shepard filter due every second step."""
    case = cases.config("small", shepard=2) if name == "shepard" else cases.config(name)
    say(f"[corner beside] case={name}")
    a = sim.open_case(case, variant=COLS)
    b = sim.open_case(case, variant=COLS)
    a.set_overlap(False)
    for graph in (True, False):
        a.step(3, graph=graph)
        b.step(3, graph=graph)
        a.sync()
        b.sync()
        for k in ("pos", "vel", "str", "cvar", "cmag", "accad", "acestn", "spin"):
            np.testing.assert_array_equal(bits(b.get(k)), bits(a.get(k)), err_msg=f"{k} graph={graph}")
    say("  3 graph + 3 eager steps bit-identical")
    a.close()
    b.close()


def test_columns_beside_timed_sections():
    """With Adami and the corner search on the side stream, their timed
    sections are recorded there, and the column index (two sections per
    call) is timed apart from the interaction kernel: eager steps and graph
    replays count the same launches per class, each class takes time, and
    the timed steps' state equals the untimed serial order's."""
    case = cases.config("small")
    say("[beside timing] case=small steps=2+4")
    a = sim.open_case(case, variant=COLS)
    b = sim.open_case(case, variant=COLS)
    ref = sim.open_case(case, variant=COLS)
    ref.set_overlap(False)
    for c in (a, b):
        c.enable_timing(True)
    a.step(2, graph=False)
    b.step(2, graph=True)
    ref.step(2, graph=True)
    a.sync()
    b.sync()
    a.enable_timing(True)  # reset the totals
    b.enable_timing(True)
    a.step(4, graph=False)
    b.step(4, graph=True)
    ref.step(4, graph=True)
    ka, kb = a.kernel_times(), b.kernel_times()
    say(f"  eager adami {ka['adami']} corner {ka['corner']}; graph adami {kb['adami']} corner {kb['corner']}")
    for k in ["neighbor_search", "sort_arrays", "adami", "corner", "column_index", "interaction", "contact",
              "predictor", "corrector"]:
        assert kb[k][1] == ka[k][1], k
        assert ka[k][0] > 0.0 and kb[k][0] > 0.0, k
    assert ka["adami"][1] == 8 and ka["corner"][1] == 8  # twice per step
    assert ka["interaction"][1] == 8 and ka["column_index"][1] == 16  # soil records, then wall records
    ref.sync()
    for k in ("pos", "vel", "str", "cvar", "cmag"):
        np.testing.assert_array_equal(bits(b.get(k)), bits(ref.get(k)), err_msg=k)
        np.testing.assert_array_equal(bits(a.get(k)), bits(ref.get(k)), err_msg=k)
    for c in (a, b, ref):
        c.close()
