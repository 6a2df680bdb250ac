"""Batched contexts (lq_create_batch): several independent wall cases stepped
by one launch per kernel must reproduce, case by case and bit for bit, what a
context holding that case alone computes -- every per-slot field after every
step, absolute positions, counts and the wall force (also with the COLUMNS
variant: a 2h column holds one case's particles, the cases being laid out
one empty cell apart in the batch grid).  Cases differ in size,
backfill height, material (phi_c, e0) and wall motion (translation / toe
rotation), and the batch is also replayed through a hipGraph."""
import sys

import numpy as np
import pytest

from retainwall_code_amd import _abi, cases, sim

pytestmark = pytest.mark.gpu

STATE = ["codes", "cellpos", "orig_id", "pos", "pos_ori", "vel", "idm", "normal", "str", "stn", "cvar",
         "cmag", "pos_ini", "vel_ini", "str_ini", "stn_ini", "accad", "spin", "acestn"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a


def small_batch(n=5, seed=3):
    return cases.batch_cases(n, seed=seed, dr=0.0025, n_range=(1500, 6000), h_range=(0.03, 0.06),
                             fix_wall_height=0.01, wall_extra=0.01)


def compare_case(batch, k, single, step):
    for name in STATE:
        a, b = batch.get_case(k, name), single.get(name)
        np.testing.assert_array_equal(bits(a), bits(b), err_msg=f"case {k} field {name} step {step}")
    np.testing.assert_array_equal(bits(batch.case_positions_absolute(k)), bits(single.positions_absolute()),
                                  err_msg=f"case {k} absolute positions step {step}")
    assert batch.case_counts(k) == single.counts()


@pytest.mark.parametrize("variant", [_abi.INTERACTION_NEIGHBOR_LIST, _abi.INTERACTION_REFERENCE_ORDER,
                                     _abi.INTERACTION_COLUMNS])
def test_batch_equals_single_cases(variant):
    cs = small_batch()
    say(f"[batch] {len(cs)} cases, np = {[c.np for c in cs]}, variant={variant}")
    batch = sim.open_batch(cs, variant=variant)
    singles = [sim.open_case(c, variant=variant) for c in cs]
    assert batch.params.np == sum(c.np for c in cs)
    for k, s in enumerate(singles):
        compare_case(batch, k, s, "initial sort")
    for step in range(3):
        batch.step(1)
        for s in singles:
            s.step(1)
        batch.sync()
        for k, s in enumerate(singles):
            compare_case(batch, k, s, step)
        say(f"  step {step}: all cases bit-identical")
    F = batch.wall_force_cases([c.probe_box for c in cs])
    for k, (c, s) in enumerate(zip(cs, singles)):
        assert F[k] == s.wall_force(c.probe_box), k
    batch.close()
    for s in singles:
        s.close()


def test_batch_graph_replay_matches_eager():
    cs = small_batch(4, seed=5)
    say(f"[batch graph] {len(cs)} cases")
    eager = sim.open_batch(cs)
    graph = sim.open_batch(cs)
    eager.step(4)
    graph.step(4, graph=True)
    eager.sync()
    graph.sync()
    for k in range(len(cs)):
        for name in ["pos", "vel", "str", "cmag", "cvar", "cellpos"]:
            np.testing.assert_array_equal(bits(eager.get_case(k, name)), bits(graph.get_case(k, name)),
                                          err_msg=f"case {k} {name}")
    eager.close()
    graph.close()


def test_batch_rejects_mismatched_physics():
    cs = small_batch(2, seed=7)
    p0 = cases.make_params(cs[0])
    p1 = cases.make_params(cs[1])
    p1.cs_m = p1.cs_m * 2.0
    with pytest.raises(sim.LqError, match="cs"):
        sim.BatchContext([p0, p1])


def mixed_model_batch():
    """Three cases with different soil models: Drucker-Prager, basic
    hypoplasticity and critical-state hypoplasticity."""
    cs = small_batch(3, seed=11)
    rows = [(_abi.DP_EPP, cases.dp_soil(phi_deg=25.0, psi_deg=5.0, cohesion=50.0)),
            (_abi.HYPO_BASIC, cases.hypo_basic_soil(phi_c_deg=32.0)), (_abi.HYPO_CS, None)]
    out = []
    for c, (mt, row) in zip(cs, rows):
        if row is None:
            out.append(c)
            continue
        c.codes[c.npb:, 2] = mt
        c.params["model_par"] = [row] + [[0.0] * 15] * 9
        out.append(c)
    return out


def test_batch_mixed_models_equal_single_cases():
    cs = mixed_model_batch()
    say(f"[batch models] {len(cs)} cases with DP, basic and critical-state soil")
    batch = sim.open_batch(cs)
    singles = [sim.open_case(c) for c in cs]
    for step in range(3):
        batch.step(1)
        for s in singles:
            s.step(1)
        batch.sync()
        for k, s in enumerate(singles):
            compare_case(batch, k, s, step)
    batch.close()
    for s in singles:
        s.close()


def test_batch_checkpoint_restart(tmp_path):
    cs = small_batch(3, seed=13)
    path = tmp_path / "batch.ckpt"
    say(f"[batch checkpoint] {len(cs)} cases")
    a = sim.open_batch(cs)
    a.step(3)
    a.checkpoint_write(path)
    a.step(4, graph=True)
    a.sync()
    ref = [{n: bits(a.get_case(k, n)) for n in STATE} for k in range(len(cs))]
    a.close()
    b = sim.open_batch(cs)
    b.checkpoint_read(path)
    b.step(4, graph=True)
    b.sync()
    for k in range(len(cs)):
        for n in STATE:
            np.testing.assert_array_equal(bits(b.get_case(k, n)), ref[k][n], err_msg=f"case {k} {n}")
    b.close()


def test_batch_with_surrogate_equals_single_cases():
    """The HypoCS surrogate (two members, guard box, 1 % audit) in a batched
    context: every case's state, and the guard/audit counts summed over the
    cases, equal those of one-case contexts with the same network.  The
    network is untrained and the guard box wide, so it answers nearly every
    call; the audit draws on (original id, step), which a case keeps in the
    batch."""
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    import mlp_ref as M
    from retainwall_code_amd import surrogate
    cs = small_batch(3, seed=11)
    say(f"[batch surrogate] {len(cs)} cases, np = {[c.np for c in cs]}")
    members = [M.init_params(s).astype(np.float32) for s in (1, 2)]
    norm = np.concatenate([np.zeros(14), np.full(14, 1e3), np.zeros(7), np.full(7, 1e-3)]).astype(np.float32)
    lo, hi = np.full(14, -1e30, np.float32), np.full(14, 1e30, np.float32)
    batch = sim.open_batch(cs)
    singles = [sim.open_case(c) for c in cs]
    for ctx in [batch] + singles:
        surrogate.install(ctx, members, norm, lo, hi, ens_tol=1e3, audit_frac=0.01)
        surrogate.enable(ctx)
    for step in range(3):
        batch.step(1)
        for s in singles:
            s.step(1)
        batch.sync()
        for k, s in enumerate(singles):
            compare_case(batch, k, s, step)
    sb = surrogate.stats(batch)
    ss = [surrogate.stats(s) for s in singles]
    say(f"  batch stats {sb}")
    for key in ("surrogate", "fallback_range", "fallback_ensemble", "audited", "tensile_branch"):
        assert sb[key] == sum(s[key] for s in ss), key
    assert sb["surrogate"] > 0 and sb["audited"] > 0
    assert sb["audit_max"] == max(s["audit_max"] for s in ss)
    batch.close()
    for s in singles:
        s.close()
