"""Periodic boundaries in the oracle (the *LoopPeriod helpers,
particle_interaction.cu:243-258, adami_boundary.cu:73-87,
field_smoothing.cu:58-72, and the corrector's re-entry,
predictor_corrector.cu:26-40).  On a lattice that is one period long in y,
the periodic images complete the neighbourhoods of the slots next to the
planes, so every field becomes invariant under lattice translations along y:
the slots at the planes must then match the interior ones, which they do
not without the periodic pass."""
import numpy as np
import pytest

from retainwall_code_amd import cases
from test_oracle import F64, soil_block


def periodic_block(oracle_mod, n, periodic, dr=0.01):
    o = soil_block(oracle_mod, n, dr=dr)
    if periodic:
        y0 = 0.02
        o.p.periodbound = 1
        o.p.planeleft[:] = [0.0, 1.0, 0.0, -y0]
        o.p.planeright[:] = [0.0, -1.0, 0.0, y0 + n[1] * dr]
    return o


def abs_pos(o):
    return o["pos"][:, :3] + o["cellpos"][:, :3] * (2 * o.p.h)


def columns(o, margin):
    """Slots at least margin from the x and z faces, grouped by (x, z)."""
    pa = abs_pos(o)
    lo, hi = pa.min(0), pa.max(0)
    keep = np.all((pa[:, [0, 2]] > lo[[0, 2]] + margin) & (pa[:, [0, 2]] < hi[[0, 2]] - margin), 1)
    key = np.round(pa[:, [0, 2]] / o.p.dr).astype(np.int64)
    return keep, key, pa


@pytest.mark.parametrize("periodic", [True, False])
def test_interaction_uniform_state_is_periodic(oracle_mod, periodic):
    o = periodic_block(oracle_mod, (10, 9, 10), periodic)
    keep, _, pa = columns(o, 2.7 * o.p.h)
    o.interaction()
    g = np.array([0.0, 0.0, np.float32(-9.81)])
    face = keep & ((pa[:, 1] < pa[:, 1].min() + 2 * o.p.h) | (pa[:, 1] > pa[:, 1].max() - 2 * o.p.h))
    assert face.sum() > 20
    err = max(np.abs(o["accad"][keep, :3] - g).max(), np.abs(o["accad"][keep, 3]).max(),
              np.abs(o["acestn"][keep]).max())
    if periodic:
        assert err < 1e-5  # the faces in y are interior points of the periodic lattice (float planes)
    else:
        assert np.abs(o["accad"][face, :3] - g).max() > 1.0  # one-sided neighbourhoods


@pytest.mark.parametrize("periodic", [True, False])
def test_shepard_of_a_periodic_field(oracle_mod, periodic):
    """A stress field cos(2 pi y / L) on the lattice: smoothing a sampled
    sinusoid over a translation-invariant neighbourhood scales it by one
    factor at every lattice row, faces included."""
    ny = 9
    o = periodic_block(oracle_mod, (10, ny, 10), periodic)
    keep, _, pa = columns(o, 2.7 * o.p.h)
    L = ny * o.p.dr
    phase = 2 * np.pi * (pa[:, 1] - 0.02) / L
    o["str"][:] = 0.0
    o["str"][:, 0] = 1000.0 * np.cos(phase)
    o["str"][:, 1] = 1000.0 * np.sin(phase)
    o.shepard()
    # smoothed = lambda (cos, sin) per slot; lambda from the interior
    lam = (o["str"][keep, 0] * np.cos(phase[keep]) + o["str"][keep, 1] * np.sin(phase[keep])) / 1000.0
    rot = (o["str"][keep, 1] * np.cos(phase[keep]) - o["str"][keep, 0] * np.sin(phase[keep])) / 1000.0
    if periodic:
        assert np.ptp(lam) < 1e-6 and np.abs(rot).max() < 1e-6 and 0.3 < lam.mean() < 1.0  # float planes
    else:
        assert np.ptp(lam) > 1e-3


def wall_case(oracle_mod, periodic):
    case = cases.retaining_wall(0.04, 0.02, 0.04, 0.0025, fix_wall_height=0.01, wall_extra=0.01,
                                periodic_y=periodic)
    o = oracle_mod.Oracle(cases.make_params(case), F64, cases.motion_tables(case)).load_host(case)
    o.abs2rel()
    o.neighbor_search()
    o.sort_arrays_primary()
    o.p.time, o.p.dt = 0.0, case.params["dt"]
    o.rigid_body_update()
    return case, o


@pytest.mark.parametrize("periodic", [True, False])
def test_adami_is_translation_invariant_along_y(oracle_mod, periodic):
    case, o = wall_case(oracle_mod, periodic)
    o.adami()
    pa = abs_pos(o)[:case.npb]
    key = np.round(pa[:, [0, 2]] / case.params["dr"]).astype(np.int64)
    h = o.p.h
    face = (pa[:, 1] < pa[:, 1].min() + 2 * h) | (pa[:, 1] > pa[:, 1].max() - 2 * h)
    spread, checked = 0.0, 0
    for k in np.unique(key, axis=0):
        m = np.all(key == k, 1)
        if m.sum() < 2:
            continue
        for f, cols in (("str", slice(0, 6)), ("vel", slice(0, 4)), ("cmag", slice(3, 6))):
            v = o[f][:case.npb][m][:, cols]
            inner = v[~face[m]]
            scale = max(np.abs(v).max(), 1e-30)
            # inside a wall block the normal integral cancels and its
            # direction is rounding noise: compare only well-defined columns
            if len(inner) and np.abs(inner - inner[0]).max() > 1e-6 * scale:
                continue
            checked += 1
            spread = max(spread, np.abs(v - v[0]).max() / scale)
    assert checked > 100
    if periodic:
        assert spread < 1e-8
    else:
        assert spread > 1e-5  # the extrapolated wall stress sees a one-sided neighbourhood


def test_corrector_reenters_through_the_other_plane(oracle_mod):
    case, o = wall_case(oracle_mod, True)
    L = case.params["planeright"][3]
    soil = np.arange(case.npb, case.np)
    pa0 = abs_pos(o)
    lowy = soil[np.argmin(pa0[soil, 1])]
    o.predictor()  # snapshots pos_ini / vel_ini
    o["vel_ini"][:, :3] = 0.0
    o["accad"][:, :3] = 0.0
    dy = -(pa0[lowy, 1] + 0.1 * case.params["dr"])  # ends 0.1 dr behind the left plane
    o["vel_ini"][lowy, 1] = dy / o.p.dt
    o["accad"][lowy, 1] = 0.0
    o.corrector()
    pa1 = abs_pos(o)
    assert abs(pa1[lowy, 1] - (pa0[lowy, 1] + dy + L)) < 1e-9
    assert 0 <= o["pos"][lowy, 1] < 2 * o.p.h
    others = np.setdiff1d(soil, [lowy])
    assert np.abs(pa1[others, 1] - pa0[others, 1]).max() < 1e-9
