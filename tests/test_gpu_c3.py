"""Full-size checks on the large sloping-backfill workload (2.0 x 1.0 x 1.0 m
at dx 4 mm, 15 deg top slope, 10 kPa strip load, translating wall):
34,024,960 particles, more than 2^25 slots, the size at which a 25-bit slot
field in a packed neighbour entry would silently wrap.

* grid invariants after a graph step (sorted keys, cell ranges, every
  particle once, counts);
* neighbour sets of 1000 sampled soil slots at or above 2^25 equal a brute
  force over all particles (k-d tree on absolute positions, then the FP32
  relative-frame predicate r^2 <= r2_cut of particle_interaction.cu:30-35);
* one interaction call with NEIGHBOR_LIST bit-identical to REFERENCE_ORDER
  (which indexes slots directly, as particle_interaction.cu:117-140), and
  both equal to the FP32 oracle, on all soil and on the slots >= 2^25;
* the COLUMNS variant: the same neighbour count for every soil slot, and
  one interaction call within the FP64 rate tolerance of the FP32 oracle.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import relerr
from retainwall_code_amd import _abi, cases, sim
from test_gpu_parity import TOL_F32_RATE, TOL_F64_RATE, bits, oracle_from, say

pytestmark = pytest.mark.gpu

BIG = 1 << 25


@pytest.fixture(scope="module")
def c3():
    case = cases.config("c3")
    assert case.np > BIG
    return case


def _grid_invariants(ctx, case):
    keys = ctx.get("part_cel")
    beg, end = ctx.get("cell_beg"), ctx.get("cell_end")
    ids = ctx.get("orig_id")
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(np.sort(ids), np.arange(case.np))
    occupied = np.unique(keys)
    np.testing.assert_array_equal(beg[occupied], np.searchsorted(keys, occupied, "left"))
    np.testing.assert_array_equal(end[occupied], np.searchsorted(keys, occupied, "right"))
    empty = np.ones(len(beg), bool)
    empty[occupied] = False
    assert np.all(beg[empty] == -1) and np.all(end[empty] == -1)
    npb, npm, ndel = ctx.counts()
    assert npb + npm + ndel == case.np
    assert np.array_equal(ctx.get("cellpos")[:, 3], keys)


def _brute_force_sets(ctx, sample, r2cut, two_h):
    """Neighbour sets (original ids) of the sampled slots over all particles,
    and the ids whose r^2 lies within a few float ulps of the cut (the
    emulated fused multiply-adds are not bit-exact there)."""
    from scipy.spatial import cKDTree
    absx = ctx.positions_absolute()[:, :3].astype(np.float64)
    rel = ctx.get("pos")[:, :3]
    cell = ctx.get("cellpos")[:, :3]
    oid = ctx.get("orig_id")
    # only particles near the sampled slots can be neighbours: those within
    # 2h (with margin) of the samples' bounding box
    lo = absx[sample].min(axis=0) - 1.1 * float(two_h)
    hi = absx[sample].max(axis=0) + 1.1 * float(two_h)
    near_box = np.nonzero(np.all((absx >= lo) & (absx <= hi), axis=1))[0]
    tree = cKDTree(absx[near_box])
    cand = tree.query_ball_point(absx[sample], r=float(two_h) * 1.05)
    sets, fuzzy = [], []
    for i, js in zip(sample, cand):
        js = near_box[np.asarray(js)]
        o = (cell[js] - cell[i]).astype(np.float32)
        d = (rel[i][None, :] - rel[js]) - o * np.float32(two_h)  # float32, pair_r2's order
        d64 = d.astype(np.float64)
        r2 = np.float32(d[:, 0] * d[:, 0]).astype(np.float64)
        r2 = (d64[:, 1] * d64[:, 1] + r2).astype(np.float32).astype(np.float64)
        r2 = (d64[:, 2] * d64[:, 2] + r2).astype(np.float32)
        keep = r2 <= r2cut
        near = np.abs(r2.astype(np.float64) - float(r2cut)) <= 4 * np.spacing(np.float32(r2cut))
        # a neighbour outside the 27 cells would be missed by every kernel
        assert np.all(np.abs(cell[js][keep] - cell[i]) <= 1)
        sets.append(set(oid[js[keep & ~near]].tolist()))
        fuzzy.append(set(oid[js[near]].tolist()))
    return sets, fuzzy


@pytest.mark.timeout(600)
def test_c3_neighbour_sets_above_2p25(c3):
    say(f"[c3 grid] np={c3.np} (2^25 = {BIG})")
    ctx = sim.open_case(c3)
    ctx.step(1, graph=True)
    ctx.sync()
    _grid_invariants(ctx, c3)
    p = ctx.get_params()
    v = C.c_float()
    assert ctx.lib.lq_interaction_cutoff_sq(C.c_float(p.i_h), C.byref(v)) == 0
    r2cut = np.float32(v.value)
    two_h = np.float32(2.0) * np.float32(p.h)
    npb, npm, _ = ctx.counts()
    say("[c3 lists] GPU neighbour counts and first ids")
    counts, first = ctx.neighbor_lists(8)
    rng = np.random.default_rng(25)
    sample = np.sort(rng.choice(np.arange(BIG, npb + npm), 1000, replace=False))
    say("[c3 lists] brute force over all particles")
    sets, fuzzy = _brute_force_sets(ctx, sample, r2cut, two_h)
    nfuzzy = 0
    for i, s, f in zip(sample, sets, fuzzy):
        k = i - npb
        g = set(first[k, :min(8, counts[k])].tolist())
        assert g - f <= s, f"slot {i}: GPU neighbour not in the brute-force set"
        assert len(s) <= counts[k] <= len(s) + len(f), f"slot {i}: count {counts[k]} vs {len(s)} (+{len(f)})"
        nfuzzy += len(f)
    say(f"  1000 slots >= 2^25: counts and ids agree ({sum(len(s) for s in sets)} pairs, {nfuzzy} at the cut)")
    ctx.set_interaction(_abi.INTERACTION_COLUMNS)
    counts_col, _ = ctx.neighbor_lists(1)
    np.testing.assert_array_equal(counts_col, counts)
    say(f"  COLUMNS windows: the same neighbour count for all {npm} soil slots")
    ctx.close()


@pytest.mark.timeout(600)
def test_c3_interaction_variants_and_oracle(c3, oracle_mod):
    ctx = sim.open_case(c3)
    ctx.step(1, graph=True)
    ctx.step_calculation()
    ctx.neighbor_search()
    ctx.sort_arrays_primary()
    ctx.rigid_body_update()
    ctx.adami()
    ctx.find_corner()
    ctx.sync()
    saved = {k: ctx.get(k) for k in ("cvar", "cmag")}
    o32 = oracle_from(ctx, oracle_mod, np.float32, c3)
    outs = {}
    for name, variant in (("nl", _abi.INTERACTION_NEIGHBOR_LIST), ("ref", _abi.INTERACTION_REFERENCE_ORDER),
                          ("col", _abi.INTERACTION_COLUMNS)):
        say(f"[c3 interaction] {name}")
        for k, v in saved.items():
            ctx.set(k, v)
        ctx.set_interaction(variant)
        ctx.interaction()
        ctx.sync()
        outs[name] = {k: ctx.get(k) for k in ("accad", "acestn", "spin", "cvar", "cmag")}
    ctx.close()
    for k in ("accad", "acestn", "spin", "cvar", "cmag"):
        np.testing.assert_array_equal(bits(outs["nl"][k]), bits(outs["ref"][k]), err_msg=k)
    say("  NEIGHBOR_LIST == REFERENCE_ORDER bit for bit")
    say("[c3 interaction] FP32 oracle (threads)")
    oracle_mod.set_surcharge(*c3.surcharge)
    try:
        o32.interaction()
    finally:
        oracle_mod.set_surcharge(-1, 0.0)
    for rows, label in ((slice(c3.npb, c3.np), "soil"), (slice(BIG, c3.np), "slots >= 2^25")):
        for k in ("accad", "acestn", "spin"):
            e = relerr(outs["nl"][k][rows], o32[k][rows])
            say(f"  {label} {k} relerr vs FP32 oracle {e:.2e}")
            assert e <= TOL_F32_RATE, (label, k)
            e = relerr(outs["col"][k][rows], o32[k][rows])
            say(f"  {label} {k} COLUMNS relerr vs FP32 oracle {e:.2e}")
            assert e <= TOL_F64_RATE, (label, k)  # other summation order and fast pair terms
        np.testing.assert_array_equal(outs["col"]["cvar"][rows][:, 2], o32["cvar"][rows][:, 2])  # slip flags


@pytest.mark.timeout(900)
def test_c3_whole_step(c3, oracle_mod):
    """One whole graph step at full size (34 M slots, above 2^25), default
    interaction (NEIGHBOR_LIST), against the FP32 and FP64 oracles' step
    from the same state: cell frames and order exact, state fields to the
    per-step tolerances of the small cases; against FP64 beyond 4h of every
    contact-wall particle (within 2h of a contact wall the contact force
    m dv / dt amplifies FP32 rounding of the velocity by 1/dt and the slip
    decisions can flip; the corrector's interaction carries those velocities
    one 2h shell further)."""
    from test_gpu_parity import TOL_F32_STATE, TOL_F64_STATE
    ctx = sim.open_case(c3)
    ctx.step(1, graph=True)
    ctx.sync()
    say("[c3 step] oracles from the GPU state")
    o32 = oracle_from(ctx, oracle_mod, np.float32, c3)
    o64 = oracle_from(ctx, oracle_mod, np.float64, c3)
    oracle_mod.set_surcharge(*c3.surcharge)
    try:
        assert o32.run(1)["ns2_identity"]
        o64.run(1)
    finally:
        oracle_mod.set_surcharge(-1, 0.0)
    ctx.step(1, graph=True)
    ctx.sync()
    from scipy.spatial import cKDTree
    g = ctx.snapshot(["pos", "vel", "str", "codes", "cellpos", "orig_id"])
    np.testing.assert_array_equal(g["cellpos"], o32["cellpos"])
    np.testing.assert_array_equal(g["orig_id"], o32["orig_id"])
    soil = slice(c3.npb, c3.np)
    pa = ctx.positions_absolute()[:, :3].astype(np.float64)
    walls = np.nonzero(g["codes"][:c3.npb, 0] >= 10)[0]
    dist, _ = cKDTree(pa[walls]).query(pa[soil], distance_upper_bound=8.0 * float(ctx.params.h))
    far = np.nonzero(dist > 4.0 * float(ctx.params.h))[0] + c3.npb
    say(f"  {len(far)} of {c3.npm} soil slots beyond 4h of every contact-wall particle")
    assert len(far) > 0.95 * c3.npm
    for k in ("pos", "vel", "str"):
        e32 = relerr(g[k][soil], o32[k][soil])
        e64 = relerr(g[k][far], o64[k][far])
        band = relerr(g[k][soil], o64[k][soil])
        say(f"  {k} relerr vs FP32 oracle {e32:.2e}; vs FP64 oracle {e64:.2e} (all soil {band:.2e})")
        assert e32 <= TOL_F32_STATE and e64 <= TOL_F64_STATE, k
    e = relerr(g["vel"][far, 3], o64["vel"][far, 3])
    say(f"  density relerr vs FP64 oracle {e:.2e}")
    assert e <= TOL_F64_STATE
    ctx.close()
