"""GPU parity for the constitutive models other than HypoCS
(constitutive_model.h:38-202, dispatched at predictor_corrector.cu:72-75 and
169-172): the wall case with Drucker-Prager, mu(I) Drucker-Prager or basic
hypoplastic soil, stepped through the C ABI and compared with the oracle's
faithful steps at the whole-step tolerances of test_gpu_parity.py."""
import numpy as np
import pytest

from retainwall_code_amd import _abi, cases, sim
from test_gpu_parity import _cmp, oracle_from, say

pytestmark = pytest.mark.gpu

MODELS = {
    # phi = 20 deg puts the K0 = 0.5 start outside the cone: every soil slot
    # returns on the first update, so the yield path (idm.x kept) is exercised
    "dp": (_abi.DP_EPP, cases.dp_soil(phi_deg=20.0, psi_deg=5.0, cohesion=0.0)),
    "mui_dp": (_abi.MUI_DP_EPP, cases.mui_dp_soil()),
    "hypo_basic": (_abi.HYPO_BASIC, cases.hypo_basic_soil()),
}


def model_case(name):
    mt, row = MODELS[name]
    c = cases.config("tiny", model_type=mt, material=row)
    # idm.x >= 100 marks soil (the cell hash, loquat_grid.cu:89-90); every
    # update resets it to 100 except a Drucker-Prager yield, which keeps it
    c.idmassidx[c.npb:, 0] = 150.0
    return c


@pytest.mark.parametrize("name", list(MODELS))
def test_model_steps(name, oracle_mod):
    case = model_case(name)
    say(f"[model steps] case=tiny model={name}")
    ctx = sim.open_case(case)
    ctx.step(2)
    ctx.sync()
    o32 = oracle_from(ctx, oracle_mod, np.float32, case)
    o64 = oracle_from(ctx, oracle_mod, np.float64, case)
    o32.run(1)
    o64.run(1)
    ctx.step(1)
    ctx.sync()
    keys = ["pos", "vel", "str", "stn", "idm", "pos_ori"]
    g = ctx.snapshot(keys + ["cellpos", "orig_id"])
    np.testing.assert_array_equal(g["cellpos"], o32["cellpos"])
    np.testing.assert_array_equal(g["orig_id"], o32["orig_id"])
    soil = slice(case.npb, case.np)
    np.testing.assert_array_equal(g["idm"][soil, 0], o32["idm"][soil, 0])
    _cmp(name, g, o32, o64, keys, soil, keys)
    if name == "dp":
        kept = np.count_nonzero(g["idm"][soil, 0] == 150.0)
        say(f"  slots whose last update yielded: {kept} of {case.npm}")
        assert kept > 0
    ctx.close()


@pytest.mark.parametrize("name", list(MODELS))
def test_model_step_pieces(name, oracle_mod):
    """Every piece of a step from one state, against both oracles; the
    predictor and the corrector carry the model-dependent stress update."""
    case = model_case(name)
    say(f"[model pieces] case=tiny model={name}")
    ctx = sim.open_case(case)
    ctx.step(2)
    ctx.sync()
    npb, np_ = case.npb, case.np
    soil, bnd = slice(npb, np_), slice(0, npb)
    ctx.step_calculation()
    ctx.neighbor_search()
    ctx.sort_arrays_primary()
    ctx.rigid_body_update()
    ctx.sync()
    seq = [("adami", bnd, ["vel", "str"], ["vel", "str"]),
           ("find_corner", soil, ["cvar", "cmag"], ["cvar", "cmag"]),
           ("interaction", soil, ["accad", "acestn", "spin"], []),
           ("predictor", soil, ["pos", "vel", "str", "stn", "idm"], ["pos", "vel", "str", "stn", "idm"]),
           ("adami", bnd, ["vel", "str"], ["vel", "str"]),
           ("find_corner", soil, ["cvar", "cmag"], ["cvar", "cmag"]),
           ("interaction", soil, ["accad", "acestn", "spin"], []),
           ("corrector", slice(0, np_), ["pos", "vel", "str", "stn", "idm", "pos_ori"],
            ["pos", "vel", "str", "stn", "idm", "pos_ori"])]
    for piece, rows, keys, state_keys in seq:
        o32 = oracle_from(ctx, oracle_mod, np.float32, case)
        o64 = oracle_from(ctx, oracle_mod, np.float64, case)
        getattr(o32, piece)()
        getattr(o64, piece)()
        getattr(ctx, piece)()
        ctx.sync()
        g = ctx.snapshot(keys)
        if "idm" in keys:
            np.testing.assert_array_equal(g["idm"][soil, 0], o32["idm"][soil, 0])
        _cmp(f"{name} {piece}", g, o32, o64, keys, rows, state_keys)
    ctx.close()
