"""HypoCS surrogate: logging, training and installation (C ABI wrappers).

The MLP (14 -> 128 -> 128 -> 128 -> 7, GELU) predicts, per soil particle and
HypoCS call (predictor_corrector.cu:75 dt/2, :172 dt), the increment
(rate - Jaumann term) * dt and M_crt from (sigma - cohesion, D dt, e, dt);
the library applies the rest of the update exactly and falls back to the
exact update outside the guard box or when the ensemble members disagree.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi

NIN, NOUT, NH, NPARAM, NNORM = _abi.MLP_IN, _abi.MLP_OUT, _abi.MLP_HIDDEN, _abi.MLP_NPARAM, _abi.MLP_NNORM


class MlpError(RuntimeError):
    pass


def _lib():
    return _abi.load_library()


def _chk(st, what):
    if st:
        raise MlpError(f"{what} failed ({st}): {_lib().lq_mlp_last_error().decode()}")


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    return a if shape is None else a.reshape(shape)


def normalisation(x, y):
    x, y = _f32(x, (-1, NIN)), _f32(y, (-1, NOUT))
    norm = np.zeros(NNORM, np.float32)
    _chk(_lib().lq_mlp_normalisation(x.ctypes.data, y.ctypes.data, len(x), norm.ctypes.data), "lq_mlp_normalisation")
    return norm


def train(x, y, norm, epochs=20, batch=1024, lr=1e-3, seed=0, params=None):
    """Adam on the GPU; returns (params, loss per epoch)."""
    x, y = _f32(x, (-1, NIN)), _f32(y, (-1, NOUT))
    init = params is None
    p = np.zeros(NPARAM, np.float32) if init else _f32(params).copy()
    hist = np.zeros(epochs, np.float32)
    _chk(_lib().lq_mlp_train(x.ctypes.data, y.ctypes.data, len(x), _f32(norm).ctypes.data, epochs, batch, lr, seed,
                             1 if init else 0, p.ctypes.data, hist.ctypes.data), "lq_mlp_train")
    return p, hist


def loss_grad(params, norm, x, y):
    x, y = _f32(x, (-1, NIN)), _f32(y, (-1, NOUT))
    g = np.zeros(NPARAM, np.float32)
    loss = C.c_double()
    _chk(_lib().lq_mlp_loss_grad(_f32(params).ctypes.data, _f32(norm).ctypes.data, x.ctypes.data, y.ctypes.data,
                                 len(x), g.ctypes.data, C.byref(loss)), "lq_mlp_loss_grad")
    return loss.value, g


def forward(params, norm, x, bf16=False):
    x = _f32(x, (-1, NIN))
    y = np.zeros((len(x), NOUT), np.float32)
    _chk(_lib().lq_mlp_forward(_f32(params).ctypes.data, _f32(norm).ctypes.data, x.ctypes.data, len(x),
                               y.ctypes.data, 1 if bf16 else 0), "lq_mlp_forward")
    return y


def guard_box(x, margin=0.05):
    """Per-feature [min, max] of the training inputs widened by margin * range."""
    x = _f32(x, (-1, NIN))
    lo, hi = x.min(0), x.max(0)
    pad = margin * (hi - lo)
    return (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)


# ------------------------------------------------------------ context side --
def install(ctx, members, norm, lo, hi, ens_tol=0.05, audit_frac=0.01):
    p = np.ascontiguousarray(np.concatenate([_f32(m) for m in members]))
    ctx._chk(ctx.lib.lq_surrogate_set(ctx.h, len(members), p.ctypes.data, _f32(norm).ctypes.data,
                                      _f32(lo).ctypes.data, _f32(hi).ctypes.data, ens_tol, audit_frac),
             "lq_surrogate_set")


def enable(ctx, on=True):
    ctx._chk(ctx.lib.lq_surrogate_enable(ctx.h, 1 if on else 0), "lq_surrogate_enable")


def stats(ctx, reset=False):
    """{'surrogate', 'fallback_range', 'fallback_ensemble', 'audited', 'tensile_branch', 'audit_max'}."""
    c = np.zeros(5, np.int64)
    am = C.c_float()
    ctx._chk(ctx.lib.lq_surrogate_stats(ctx.h, c.ctypes.data, C.byref(am), 1 if reset else 0), "lq_surrogate_stats")
    keys = ["surrogate", "fallback_range", "fallback_ensemble", "audited", "tensile_branch"]
    out = {k: int(v) for k, v in zip(keys, c)}
    out["audit_max"] = am.value
    return out


def start_log(ctx, every=8, capacity=1_000_000):
    ctx._chk(ctx.lib.lq_surrogate_log(ctx.h, every, capacity), "lq_surrogate_log")


def read_log(ctx, max_rows=None, sort_keys=None):
    """Logged (x, y) rows, in lexicographic order of (x, y).

    Waves reserve log rows with an atomic counter, so the device order of the
    rows depends on scheduling; sorting makes the returned data set (and what
    is trained on it) reproducible whenever the log did not overflow."""
    cap = max_rows or 1_000_000
    x = np.zeros((cap, NIN), np.float32)
    y = np.zeros((cap, NOUT), np.float32)
    n = C.c_int()
    ctx._chk(ctx.lib.lq_surrogate_log_read(ctx.h, x.ctypes.data, y.ctypes.data, cap, C.byref(n)),
             "lq_surrogate_log_read")
    x, y = x[:n.value], y[:n.value]
    if sort_keys is None:
        order = np.lexsort(np.concatenate([x, y], axis=1).T[::-1])
    else:  # a few input columns: cheaper, and as reproducible unless rows tie in all of them
        order = np.lexsort(x[:, list(sort_keys)[::-1]].T)
    return x[order].copy(), y[order].copy()


def collect(case, steps, rows, graph=True):
    """Exact steps of a fresh context of `case`, logging about `rows` (x, y)
    tuples spread over the whole run: every k-th soil slot per HypoCS call
    with k = 2 * steps * npm / rows.  Returns the rows (sorted on four input
    columns)."""
    from . import sim
    ctx = sim.open_case(case)
    every = max(1, (2 * steps * case.npm) // max(rows, 1))
    cap = (2 * steps * case.npm) // every + case.npm // every + 1024
    start_log(ctx, every=every, capacity=cap)
    ctx.step(steps, graph=graph)
    ctx.sync()
    x, y = read_log(ctx, max_rows=cap, sort_keys=(0, 1, 2, 6))
    ctx.close()
    return x, y
