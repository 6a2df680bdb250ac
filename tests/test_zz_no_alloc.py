"""lq_wall_pressure_profile_noalloc against lq_wall_pressure_profile: the
same rows bit for bit, on two calls in a row (a stepped 30k-particle case
with 32 bins, then a 6k-particle case with 12 bins, so the second call leaves
the first call's larger contents in the shared scratch), and sizes beyond
LQ_PROFILE_NOALLOC_MAX_BINS are rejected."""
import ctypes as C

import numpy as np
import pytest

from retainwall_code_amd import _abi, cases, sim

pytestmark = pytest.mark.gpu


def _geom(ctx, height):
    p = ctx.get_params()
    bt = np.float32(p.bound_thickness)
    fwh = np.float32(0.01)
    t, _ = ctx.time()
    return (np.float32(fwh + bt - np.float32(t) * np.float32(0.001)), bt + fwh, bt + fwh + np.float32(height),
            bt + fwh + np.float32(height))


def test_noalloc_profile_matches_allocating_entry_point():
    for name, height, nsteps in (("small", 0.08, 6), ("tiny", 0.03, 4)):
        case = cases.config(name)
        ctx = sim.open_case(case)
        ctx.step(nsteps)
        ctx.sync()
        geom = _geom(ctx, height)
        nbins = int(np.float32(height) / np.float32(case.params["dr"]))
        print(f"[noalloc] case={name} nbins={nbins}", flush=True)
        want = ctx.wall_pressure_profile(geom, nbins)
        got = ctx.wall_pressure_profile_noalloc(geom, nbins)
        np.testing.assert_array_equal(got.view(np.int32), want.view(np.int32))
        assert np.abs(want[:, 1]).max() > 0
        ctx.close()


def test_noalloc_profile_rejects_too_many_bins():
    case = cases.config("tiny")
    ctx = sim.open_case(case)
    n = _abi.PROFILE_NOALLOC_MAX_BINS + 1
    out = np.zeros((n, 4), np.float32)
    g = (C.c_float * 4)(0.0, 0.0, 1.0, 1.0)
    st = ctx.lib.lq_wall_pressure_profile_noalloc(ctx.h, g, n, out.ctypes.data, out.nbytes)
    assert st == _abi.ERR_ARG
    assert b"LQ_PROFILE_NOALLOC_MAX_BINS" in ctx.lib.lq_last_error(ctx.h)
    ctx.close()
