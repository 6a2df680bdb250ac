"""HypoCS surrogate on the GPU.

* FP32 training kernels: loss and gradient equal the FP64 reference
  (tests/mlp_ref.py, itself pinned by finite differences) within FP32
  rounding; the FP32 forward likewise; the BF16-MFMA forward within BF16
  rounding of weights and activations.
* Training reduces the loss.
* Logged (x, y) tuples are the exact update's: M_crt and the stress
  increment agree with the FP64 oracle's HypoCS for the same state.
* Guard: with an empty guard box every call falls back to the exact update
  and whole steps are bit-identical to runs without the surrogate.
* End to end: trained members accept most calls, about 1 % is audited, and
  stresses stay close to the exact run.
"""
import sys

import numpy as np
import pytest

import mlp_ref as M
from conftest import relerr
from retainwall_code_amd import cases, sim, surrogate

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed).astype(np.float32)
    norm = np.concatenate([rng.normal(0, 1, 14), rng.uniform(0.5, 2, 14), rng.normal(0, 1, 7),
                           rng.uniform(0.5, 2, 7)]).astype(np.float32)
    x = (rng.normal(0, 1, (n, 14)) * norm[14:28] + norm[:14]).astype(np.float32)
    t = (rng.normal(0, 1, (n, 7)) * norm[35:42] + norm[28:35]).astype(np.float32)
    return p, norm, x, t


def test_mlp_gradient_matches_fp64_reference():
    p, norm, x, t = random_problem(3000, 11)
    say("[mlp] gradient, 3000 rows")
    loss, g = surrogate.loss_grad(p, norm, x, t)
    lr, gr = M.loss_grad(p.astype(np.float64), norm.astype(np.float64), x, t)
    assert abs(loss - lr) <= 1e-5 * lr
    assert relerr(g, gr, 1e-30) < 1e-4


def test_mlp_forward_fp32_and_bf16():
    p, norm, x, _ = random_problem(5000, 12)
    say("[mlp] forward, 5000 rows")
    ref = M.forward(p.astype(np.float64), norm.astype(np.float64), x)
    y32 = surrogate.forward(p, norm, x, bf16=False)
    assert relerr(y32, ref, 1e-30) < 1e-5
    y16 = surrogate.forward(p, norm, x, bf16=True)
    # weights and activations rounded to bf16 (8 significant bits) over 4 layers
    e16 = relerr(y16, ref, 1e-30)
    say(f"  bf16 normwise relative error {e16:.3e}")
    assert e16 < 3e-2
    # rows beyond a partial 16-row tile and a single row
    y1 = surrogate.forward(p, norm, x[:1], bf16=True)
    np.testing.assert_array_equal(y1, y16[:1])
    y17 = surrogate.forward(p, norm, x[:17], bf16=True)
    np.testing.assert_array_equal(y17, y16[:17])


def test_training_reduces_loss():
    rng = np.random.default_rng(13)
    x = rng.uniform(-1, 1, (20000, 14)).astype(np.float32)
    w = rng.normal(0, 1, (14, 7))
    t = (np.tanh(x @ w) + 0.1 * x[:, :7] ** 2).astype(np.float32)
    norm = surrogate.normalisation(x, t)
    say("[mlp] training, 20000 rows x 8 epochs")
    p, hist = surrogate.train(x, t, norm, epochs=8, batch=512, lr=2e-3, seed=1)
    say(f"  loss per epoch {np.round(hist, 5)}")
    assert hist[-1] < 0.1 * hist[0]
    l_end, _ = M.loss_grad(p.astype(np.float64), norm.astype(np.float64), x, t)
    assert l_end < 0.1 * hist[0]


def test_log_matches_exact_update(oracle_mod):
    case = cases.config("small")
    say("[surrogate] log, case=small, 2 steps")
    ctx = sim.open_case(case)
    surrogate.start_log(ctx, every=4, capacity=200000)
    ctx.step(2)
    ctx.sync()
    x, y = surrogate.read_log(ctx)
    say(f"  logged {len(x)} tuples")
    assert len(x) > 1000
    mp = np.array(case.params["model_par"][0], np.float64)
    rng = np.random.default_rng(0)
    rows = rng.choice(len(x), size=300, replace=False)
    close = 0
    for r in rows:
        s, ddt, e, dt = x[r, :6].astype(np.float64), x[r, 6:12].astype(np.float64), float(x[r, 12]), float(x[r, 13])
        rho = (1.0 + mp[11]) * mp[12] / (1.0 + e)
        sig, mcrt = oracle_mod.hypo_cs(dt, rho, np.zeros(3), ddt / dt, s.copy(), mp, np.float64)
        assert abs(mcrt - y[r, 6]) <= 1e-4 * max(1.0, abs(mcrt)), (r, mcrt, y[r, 6])
        inc = np.asarray(sig) - s
        if np.max(np.abs(inc - y[r, :6])) <= 1e-3 * np.max(np.abs(inc)) + 1e-6 * np.max(np.abs(s)):
            close += 1
    say(f"  stress increment within tolerance for {close} of {len(rows)} rows")
    assert close >= 0.9 * len(rows)  # the rest: the oracle's shear cap rescaled the stress
    ctx.close()


def test_guard_fallback_is_exact():
    case = cases.config("small")
    say("[surrogate] empty guard box, case=small, 3 steps")
    plain = sim.open_case(case)
    sur = sim.open_case(case)
    p = M.init_params(1).astype(np.float32)
    norm = np.concatenate([np.zeros(14), np.ones(14), np.zeros(7), np.ones(7)]).astype(np.float32)
    lo = np.full(14, 1.0, np.float32)
    hi = np.full(14, -1.0, np.float32)  # lo > hi: nothing is in range
    surrogate.install(sur, [p, p], norm, lo, hi, ens_tol=1.0, audit_frac=0.01)
    surrogate.enable(sur)
    plain.step(3)
    sur.step(3)
    plain.sync()
    sur.sync()
    st = surrogate.stats(sur)
    say(f"  stats {st}")
    assert st["surrogate"] == 0 and st["fallback_range"] > 0 and st["audited"] == 0
    for k in ["str", "idm", "vel", "pos", "cvar", "cmag", "stn"]:
        np.testing.assert_array_equal(bits(plain.get(k)), bits(sur.get(k)), err_msg=k)
    plain.close()
    sur.close()


def test_surrogate_end_to_end():
    case = cases.config("small")
    say("[surrogate] train on logged tuples, case=small")
    ctx = sim.open_case(case)
    surrogate.start_log(ctx, every=1, capacity=400000)
    ctx.step(6)
    ctx.sync()
    x, y = surrogate.read_log(ctx)
    ctx.close()
    say(f"  {len(x)} tuples")
    norm = surrogate.normalisation(x, y)
    members = [surrogate.train(x, y, norm, epochs=6, batch=1024, lr=1e-3, seed=s)[0] for s in (1, 2)]
    lo, hi = surrogate.guard_box(x)
    plain = sim.open_case(case)
    sur = sim.open_case(case)
    plain.step(6)
    sur.step(6)
    surrogate.install(sur, members, norm, lo, hi, ens_tol=0.25, audit_frac=0.01)
    surrogate.enable(sur)
    plain.step(4)
    sur.step(4)
    plain.sync()
    sur.sync()
    st = surrogate.stats(sur)
    say(f"  stats {st}")
    total = st["surrogate"] + st["fallback_range"] + st["fallback_ensemble"]
    assert st["surrogate"] > 0.5 * total
    assert 0.002 * st["surrogate"] <= st["audited"] <= 0.03 * st["surrogate"]
    soil = slice(case.npb, case.np)
    e = relerr(sur.get("str")[soil], plain.get("str")[soil], 1e-3)
    say(f"  soil stress normwise relative difference vs exact after 4 steps: {e:.3e}")
    assert e < 1e-2
    plain.close()
    sur.close()
