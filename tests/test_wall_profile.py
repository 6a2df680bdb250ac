"""Wall pressure profile (cuGetValuesAlongLine, utilities.cu:266-320): the
oracle's serial restatement against a vectorised statement of the same
rule (the largest slot index in each band and depth bin is the serial
loop's last writer), band edges counted in both bands, slots outside the
depth range ignored; and the GPU entry point bit-exact against the oracle
on a stepped wall case."""
import numpy as np
import pytest

from retainwall_code_amd import cases
from test_oracle import F32


def expected(o, geom, nbins):
    """Vectorised rule: per (band, bin) the max slot index wins."""
    p = o.p
    pa = o["pos"][:, :3] + o["cellpos"][:, :3].astype(np.float32) * np.float32(2 * np.float32(p.h))
    pa = pa.astype(np.float32)
    dr = np.float32(p.dr)
    x0, zlo, zhi, ztop = (np.float32(g) for g in geom)
    idx = np.arange(p.npb, p.np)
    z = pa[idx, 2]
    x = pa[idx, 0]
    inz = (z >= zlo) & (z <= zhi)
    b = ((ztop - z) / dr).astype(np.int64)
    out = np.zeros((nbins, 4))
    for band in range(3):
        lo = np.float32(x0 + dr * np.float32(band))
        hi = np.float32(x0 + dr * np.float32(band + 1))
        m = inz & (x >= lo) & (x <= hi) & (b >= 0) & (b < nbins)
        for k in np.unique(b[m]):
            last = idx[m & (b == k)].max()
            out[k, 1 + band] = np.float32(o["cmag"][last, 0] / dr)
            if band == 0:
                out[k, 0] = np.float32(np.float32(k) * dr)
    return out


def wall_oracle(oracle_mod):
    case = cases.config("tiny")
    o = oracle_mod.Oracle(cases.make_params(case), F32, cases.motion_tables(case)).load_host(case)
    o.abs2rel()
    o.neighbor_search()
    o.sort_arrays_primary()
    o["cmag"][:, 0] = np.arange(case.np, dtype=np.float32) + 1.0  # distinct force per slot
    return case, o


def test_profile_matches_last_writer_rule(oracle_mod):
    case, o = wall_oracle(oracle_mod)
    p = o.p
    bt, fwh, dr = np.float32(p.bound_thickness), np.float32(0.01), np.float32(p.dr)
    geom = (fwh + bt, bt + fwh, bt + fwh + np.float32(0.05), bt + fwh + np.float32(0.05))
    nbins = int(np.float32(0.05) / dr)
    got = o.wall_profile(geom, nbins)
    want = expected(o, geom, nbins)
    np.testing.assert_array_equal(got, want)
    assert np.count_nonzero(got[:, 1]) > nbins // 2  # the first band is populated


def test_profile_band_edges_and_range(oracle_mod):
    case, o = wall_oracle(oracle_mod)
    p = o.p
    dr = np.float32(p.dr)
    s = case.npb + 5  # put one soil slot exactly on the band-1/band-2 edge
    pa = o["pos"][s, :3] + o["cellpos"][s, :3] * np.float32(2 * np.float32(p.h))
    x_face = np.float32(pa[0] - dr)
    geom = (x_face, np.float32(pa[2] - dr), np.float32(pa[2] + dr), np.float32(pa[2] + dr))
    got = o.wall_profile(geom, 4)
    want = expected(o, geom, 4)
    np.testing.assert_array_equal(got, want)
    k = int((np.float32(pa[2] + dr) - np.float32(pa[2])) / dr)
    if float(x_face + dr) == float(pa[0]):  # exactly on the edge: in both bands
        assert got[k, 1] != 0 and got[k, 2] != 0
    # nothing outside [z_lo, z_hi]
    none = o.wall_profile((x_face, np.float32(1e3), np.float32(2e3), np.float32(2e3)), 4)
    assert not none.any()


@pytest.mark.gpu
def test_gpu_profile_bitexact(oracle_mod):
    from retainwall_code_amd import sim
    from test_gpu_parity import oracle_from, say
    case = cases.config("small")
    say("[profile] case=small, 6 steps")
    ctx = sim.open_case(case)
    ctx.step(6)
    ctx.sync()
    p = ctx.get_params()
    bt, dr = np.float32(p.bound_thickness), np.float32(p.dr)
    fwh = np.float32(0.01)
    t, _ = ctx.time()
    geom = (np.float32(fwh + bt - np.float32(t) * np.float32(0.001)), bt + fwh, bt + fwh + np.float32(0.08),
            bt + fwh + np.float32(0.08))
    nbins = int(np.float32(0.08) / dr)
    g = ctx.wall_pressure_profile(geom, nbins)
    o = oracle_from(ctx, oracle_mod, np.float32, case)
    ref = o.wall_profile(geom, nbins)
    say(f"  bins with band-1 slots: {np.count_nonzero(g[:, 0])} of {nbins}; max |p1| {np.abs(g[:, 1]).max():.4g}")
    np.testing.assert_array_equal(g, ref.astype(np.float32))
    assert np.abs(g[:, 1]).max() > 0  # contact forces present
    ctx.close()
