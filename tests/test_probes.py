"""Pins of the oracle's point probes (no GPU): the stress sensor of
GetSoilStress (utilities.cu:470-528) and the surface heights of GetSurface /
GetSurfaceAtPoint (utilities.cu:26-91, 371-427), restated with the points
as arguments.

* the sensor preserves a constant stress field, reproduces a linear field at
  a lattice particle (symmetric neighbourhood: the first moment of W
  vanishes), sees exactly the brute-force set of soil particles within 2h
  (all 27 cells) or within 2h and in the point's own y row of cells (the
  reference's cell set), and its weight sum is the lattice sum of W V;
* surface heights on a lattice block equal the top layer's z on lattice
  columns, find nothing midway between columns at radius 0.5 dr (strict
  test), and on a jittered block equal a brute-force maximum over the
  particles inside the disc or the strip.
"""
import math

import numpy as np

from test_oracle import F32, F64, soil_block


def _abs(o):
    return (o["pos"][:, :3] + o["cellpos"][:, :3] * (2 * o.p.h)).astype(np.float64)


def _wendland(knorm, q):
    t = 2.0 - q
    return knorm * t ** 4 * (q + 0.5)  # 3D Wendland C2, h-scaled (functions.h:142)


def test_sensor_constant_linear_and_sets(oracle_mod):
    o = soil_block(oracle_mod, (14, 12, 12), dtype=F64)
    pa = _abs(o)
    h = float(o.p.h)
    # lattice particles well inside the block
    lo, hi = pa.min(0) + 2.5 * h, pa.max(0) - 2.5 * h
    inner = np.nonzero(np.all((pa > lo) & (pa < hi), 1))[0][::37][:12]
    pts = pa[inner].astype(np.float32)
    sig0 = np.array([-1000.0, -1000.0, -2000.0, 0.0, 0.0, 0.0])
    out = o.sensor_stress(pts)
    np.testing.assert_allclose(out[:, :6], np.broadcast_to(sig0, (len(pts), 6)), rtol=1e-12, atol=1e-9)
    # linear field sigma_xx = a + b x: Shepard at a lattice particle returns its own value (to the
    # float32 rounding of the point)
    a, b = -500.0, 2.0e4
    o["str"][:, 0] = a + b * pa[:, 0]
    out = o.sensor_stress(pts)
    np.testing.assert_allclose(out[:, 0], a + b * pa[inner, 0], rtol=1e-7)  # points rounded to float32
    # brute-force sets and weight sums, both cell sets
    V = o["idm"][:, 1] / o["vel"][:, 3]
    for rows in (False, True):
        out = o.sensor_stress(pts, reference_rows=rows)
        for k, x in enumerate(pts.astype(np.float64)):
            r = np.linalg.norm(pa - x, axis=1)
            sel = r / h < 2.0
            if rows:
                cy = math.floor(0.5 * float(x[1]) * float(o.p.i_h))
                sel &= o["cellpos"][:, 1] == cy
            assert int(out[k, 7]) == int(sel.sum())
            w = np.sum(_wendland(float(o.p.kernel_normalization_par), r[sel] / h) * V[sel])
            assert abs(out[k, 6] - w) <= 1e-7 * w  # float32 kernel constant here
        if not rows:  # partition of unity on an interior lattice point, up to the kernel's lattice error
            assert np.all(np.abs(out[:, 6] - 1.0) < 0.05)
            full = out[:, 7].copy()
        else:  # one row of cells sees a subset, a proper one for some points
            assert np.all(out[:, 7] <= full) and np.any(out[:, 7] < full)


def test_sensor_far_point_and_fp32(oracle_mod):
    o64 = soil_block(oracle_mod, (8, 8, 8), dtype=F64)
    o32 = soil_block(oracle_mod, (8, 8, 8), dtype=F32)
    far = np.array([[0.9, 0.9, 0.9]], np.float32)  # outside the grid's particles
    out = o64.sensor_stress(far)
    assert out[0, 7] == 0 and out[0, 6] == 0.0 and np.all(out[0, :6] == 0.0)
    pts = _abs(o64)[::53][:8].astype(np.float32)
    a, b = o32.sensor_stress(pts), o64.sensor_stress(pts)
    assert np.all(a[:, 7] == b[:, 7])
    np.testing.assert_allclose(a[:, :7], b[:, :7], rtol=2e-6, atol=1e-3)


def test_surface_heights_lattice_and_brute_force(oracle_mod):
    o = soil_block(oracle_mod, (12, 10, 9), dtype=F64)
    pa = _abs(o)
    dr, h = float(o.p.dr), float(o.p.h)
    ztop = pa[:, 2].max()
    cols = np.unique(np.round(pa[:, :2], 9), axis=0)
    pick = cols[::7]
    out = o.surface_heights(pick.astype(np.float32), 0.5 * dr)
    np.testing.assert_allclose(out, ztop, rtol=0, atol=1e-9)
    # midway between four columns: every column is 0.5 dr * sqrt(2) away;
    # midway between two columns along x: exactly 0.5 dr away -> strict test fails
    mid = (pick[:3] + 0.5 * dr).astype(np.float32)
    assert np.all(o.surface_heights(mid, 0.5 * dr) == -1.0e10)
    # strip rule: |dx| < 2h over the point's own row of cells
    out = o.surface_heights(pick.astype(np.float32), -2.0 * h)
    np.testing.assert_allclose(out, ztop, rtol=0, atol=1e-9)
    # jittered block with a sloped top: brute force over the disc / strip
    oj = soil_block(oracle_mod, (12, 10, 9), dtype=F64, jitter=0.2, seed=3)
    pj = _abs(oj)
    keep = pj[:, 2] < 0.06 + 0.3 * (pj[:, 0] - 0.02)  # drop particles above a slope
    oj["cellpos"][~keep, 2] = -1000  # far below: never a top (cell ranges unchanged)
    pj = _abs(oj)
    rng = np.random.default_rng(1)
    q = np.stack([rng.uniform(0.03, 0.13, 20), rng.uniform(0.03, 0.11, 20)], 1).astype(np.float32)
    for radius in (0.5 * dr, 1.3 * dr, -2.0 * h):
        got = oj.surface_heights(q, radius)
        for k, (x, y) in enumerate(q.astype(np.float64)):
            if radius > 0:
                sel = np.hypot(pj[:, 0] - x, pj[:, 1] - y) < radius
            else:
                cy = math.floor(0.5 * y * float(oj.p.i_h))
                sel = (np.abs(pj[:, 0] - x) < -radius) & (oj["cellpos"][:, 1] == cy)
            want = pj[sel, 2].max() if sel.any() else -1.0e10
            assert abs(got[k] - want) <= 1e-9 * max(1.0, abs(want)), (radius, k)
