"""Morton cell keys (LQ_KEY_MORTON).

* the keys are the interleaved cell coordinates (checked against a bit-by-bit
  numpy interleave), sorted stably, with consistent cell ranges;
* per particle (matched by original id) the neighbour sets equal those of
  linear keys;
* steps with Morton keys reproduce the steps with linear keys per particle
  (bit for bit while no particle changes cell, within FP32 rounding after),
  on the small case and at full c2 size.
"""
import numpy as np
import pytest

from conftest import relerr
from retainwall_code_amd import _abi, cases, sim
from test_gpu_parity import TOL_F32_STATE, say, shuffled

pytestmark = pytest.mark.gpu
NL = _abi.INTERACTION_NEIGHBOR_LIST


def morton_np(c):
    """Interleave the bits of (x, y, z): x above y above z, z lowest."""
    key = np.zeros(len(c), np.int64)
    for b in range(10):
        for a, sh in ((0, 2), (1, 1), (2, 0)):
            key |= ((c[:, a].astype(np.int64) >> b) & 1) << (3 * b + sh)
    return key


def by_id(ctx, name):
    a = ctx.get(name)
    out = np.empty_like(a)
    out[ctx.get("orig_id")] = a
    return out


@pytest.mark.parametrize("name,shuffle", [("tiny", True), ("small", False)])
def test_morton_keys_and_ranges(name, shuffle):
    case = cases.config(name)
    if shuffle:
        case = shuffled(case, 9)
    ctx = sim.open_case(case, key_mode=_abi.KEY_MORTON, variant=NL)
    p = ctx.get_params()
    b = 0
    while (1 << b) < max(p.cell_num[0], p.cell_num[1], p.cell_num[2]):
        b += 1
    assert p.ncell_material == 1 << (3 * b) and p.ncell == 2 * p.ncell_material + 1
    keys = ctx.get("part_cel")
    cp = ctx.get("cellpos")
    want = morton_np(cp[:, :3]) + np.where(np.arange(case.np) >= case.npb, p.ncell_material, 0)
    np.testing.assert_array_equal(keys, want)
    np.testing.assert_array_equal(cp[:, 3], keys)
    assert np.all(np.diff(keys) >= 0)
    # stable: inside a cell the upload order (original id) is kept at the first sort
    ids = ctx.get("orig_id")
    same = np.diff(keys) == 0
    assert np.all(np.diff(ids)[same] > 0)
    beg, end = ctx.get("cell_beg"), ctx.get("cell_end")
    occ = np.unique(keys)
    np.testing.assert_array_equal(beg[occ], np.searchsorted(keys, occ, "left"))
    np.testing.assert_array_equal(end[occ], np.searchsorted(keys, occ, "right"))
    empty = np.ones(len(beg), bool)
    empty[occ] = False
    assert np.all(beg[empty] == -1)
    say(f"[morton keys] {case.name}: {len(occ)} occupied of {p.ncell} keys")
    ctx.close()


def test_morton_neighbor_sets_and_steps():
    case = cases.config("small")
    a = sim.open_case(case, key_mode=_abi.KEY_LINEAR, variant=NL)
    b = sim.open_case(case, key_mode=_abi.KEY_MORTON, variant=NL)
    for n in (0, 3, 6):
        if n:
            a.step(3)
            b.step(3, graph=True)
            a.sync()
            b.sync()
        ca, ia = a.neighbor_lists(200)
        cb, ib = b.neighbor_lists(200)
        oa = a.get("orig_id")[case.npb:] - case.npb
        ob = b.get("orig_id")[case.npb:] - case.npb
        sa = np.empty_like(ca); sa[oa] = ca
        sb = np.empty_like(cb); sb[ob] = cb
        np.testing.assert_array_equal(sa, sb)
        la = {int(o): sorted(ia[k, :ca[k]].tolist()) for k, o in enumerate(oa)}
        lb = {int(o): sorted(ib[k, :cb[k]].tolist()) for k, o in enumerate(ob)}
        assert la == lb
        bitsame = True
        for k in ("pos", "vel", "str", "cvar", "cmag"):
            x, y = by_id(a, k), by_id(b, k)
            bitsame &= np.array_equal(x.view(np.int32), y.view(np.int32))
            assert relerr(y[case.npb:], x[case.npb:]) <= TOL_F32_STATE, k
        say(f"[morton steps] after {n} steps: sets equal, fields {'bit-identical' if bitsame else 'within FP32'}")
    a.close()
    b.close()


@pytest.mark.timeout(300)
def test_morton_c2_steps():
    c2 = cases.config("c2")
    a = sim.open_case(c2, key_mode=_abi.KEY_LINEAR, variant=NL)
    b = sim.open_case(c2, key_mode=_abi.KEY_MORTON, variant=NL)
    a.step(4, graph=True)
    b.step(4, graph=True)
    a.sync()
    b.sync()
    for k in ("pos", "vel", "str"):
        x, y = by_id(a, k), by_id(b, k)
        e = relerr(y[c2.npb:], x[c2.npb:])
        say(f"[morton c2] {k} relerr Morton vs linear {e:.2e}")
        assert e <= TOL_F32_STATE, k
    a.close()
    b.close()
