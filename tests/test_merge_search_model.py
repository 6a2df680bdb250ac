"""The index formulas of the merge search (k_mover_compact, k_merge_place in
retainwall_code_amd/csrc/lq_grid.hip), restated in numpy and checked against
a stable sort by (key, slot), the order the reference's stable radix sort
gives (loquat_grid.cu:98-113).  After a previous search slot i held key
prev[i] with prev sorted; slots whose key is unchanged keep their order, the
movers are sorted stably by key, and each element's position is its rank in
its own sequence plus a count from the other.  The GPU kernels themselves
are checked in tests/test_gpu_parity.py (test_merge_search_many_movers and
every stepping test)."""
import numpy as np
import pytest


def merge_positions(keys, prev):
    n = len(keys)
    flag = (keys != prev).astype(np.int64)
    mscan = np.concatenate([[0], np.cumsum(flag)])  # exclusive scan, count at [n]
    M = int(mscan[n])
    movers = np.nonzero(flag)[0]  # slot order, as compacted
    order = np.argsort(keys[movers], kind="stable")
    smkey, smslot = keys[movers][order], movers[order]
    out_key = np.full(n, -1)
    out_idx = np.full(n, -1)
    for i in range(n):
        if mscan[i + 1] == mscan[i]:  # stayer
            key, lo, hi = keys[i], 0, M
            while lo < hi:
                mid = (lo + hi) >> 1
                if smkey[mid] < key or (smkey[mid] == key and smslot[mid] < i):
                    lo = mid + 1
                else:
                    hi = mid
            pos = (i - mscan[i]) + lo
            assert out_idx[pos] == -1
            out_key[pos], out_idx[pos] = key, i
        if i < M:  # the i-th sorted mover
            key, slot = smkey[i], smslot[i]
            lo, hi = 0, n
            while lo < hi:
                mid = (lo + hi) >> 1
                if prev[mid] < key:
                    lo = mid + 1
                else:
                    hi = mid
            up, uh = lo, n
            while up < uh:
                mid = (up + uh) >> 1
                if prev[mid] <= key:
                    up = mid + 1
                else:
                    uh = mid
            p = max(lo, min(slot, up))
            pos = i + (p - mscan[p])
            assert out_idx[pos] == -1
            out_key[pos], out_idx[pos] = key, slot
    return out_key, out_idx


@pytest.mark.parametrize("mode", ["few", "many", "none", "all_and_deleted"])
def test_merge_equals_stable_sort(mode):
    rng = np.random.default_rng({"few": 1, "many": 2, "none": 3, "all_and_deleted": 4}[mode])
    for _ in range(100):
        n = int(rng.integers(1, 250))
        nk = int(rng.integers(1, 40))
        prev = np.sort(rng.integers(0, nk, n))
        keys = prev.copy()
        moved = {"few": rng.random(n) < 0.05, "many": rng.random(n) < 0.7, "none": np.zeros(n, bool),
                 "all_and_deleted": np.ones(n, bool)}[mode]
        keys[moved] = rng.integers(0, nk + 1, int(moved.sum()))
        if mode == "all_and_deleted":
            keys[rng.random(n) < 0.1] = nk + 5  # the deleted-particle key sorts last
        ok, oi = merge_positions(keys, prev)
        ref = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(oi, ref)
        np.testing.assert_array_equal(ok, keys[ref])
