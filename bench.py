"""Benchmark: plain SPH retaining-wall step on one MI355X.

Workload: the production-size wall case (1.0 x 0.5 x 0.6 m backfill at
dx = 0.0063 m, ~1.38M soil particles, toe rotation, jittered lattice and
random void ratio), FP32.  A step is one predictor-corrector time step of the
whole case (neighbour search, sort, Adami wall, corner detection, two
interaction passes with wall contact, hypoplastic updates).

Prints one JSON line with particle-steps/s, the interaction kernel's HBM
roofline (algorithmic bytes / measured launch time) and the CPU baseline
(the FP32 oracle, serial, on a bounded sample of the same case).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3     # MI355X vector FP32
# Compulsory bytes per interaction call, each field once: 168 B per soil
# particle (in: position 12, velocity 12, density 4, mass 4, stress 24,
# flags 4, contact state 36; out: acceleration 12, drho/dt 4, D 24, spin 12,
# contact force 12, slip 4, K 4), 56 B per wall particle (position,
# velocity, density, stress, flags) and 8 B per cell of the linear grid
# (begin/end), whatever key function the context uses.
B_SOIL, B_BND, B_CELL = 168, 56, 8
# Flops of the interaction kernel counted from its source (FMA = 2):
# 15 per candidate distance test, 110 per neighbour pair evaluation.
FLOP_CAND, FLOP_PAIR = 15, 110


def _morton(x, y, z):
    key = np.zeros(len(x), np.int64)
    for b in range(10):
        key |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
    return key


def candidate_counts(cellpos, beg, end, p, soil):
    """Per soil particle: candidates scanned (27 cells x 2 ranges)."""
    cnt = np.where(beg >= 0, end - beg, 0).astype(np.int64)
    nx, ny, nz = p.cell_num
    ncm = p.ncell_material
    c = cellpos[soil, :3].astype(np.int64)
    tot = np.zeros(len(c), np.int64)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x, y, z = c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz
                ok = (x >= 0) & (y >= 0) & (z >= 0) & (x < nx) & (y < ny) & (z < nz)
                if p.key_mode == 1:
                    k = np.where(ok, _morton(np.maximum(x, 0), np.maximum(y, 0), np.maximum(z, 0)), 0)
                else:
                    k = np.where(ok, x * ny * nz + y * nz + z, 0)
                tot += np.where(ok, cnt[k] + cnt[k + ncm], 0)
    return tot


def col_candidate_counts(pos, cellpos, p, soil, np_live):
    """Per soil particle: candidates the COLUMNS variant tests (its sub-column
    windows, with the kernel's margins), from the same float32 state."""
    h, i_h = np.float32(p.h), np.float32(p.i_h)
    nx, ny, nz = p.cell_num
    rel = pos[:np_live, :3]
    cp = cellpos[:np_live, :3].astype(np.int64)
    sub = (rel >= h).astype(np.int64)
    xs = 2 * cp + sub
    ncol2 = 2 * nz
    bucket = ((xs[:, 0] * (2 * ny) + xs[:, 1]) * ncol2 + xs[:, 2])
    nb = (2 * nx) * (2 * ny) * ncol2
    pre = np.concatenate([[0], np.cumsum(np.bincount(bucket, minlength=nb))])
    f = (rel - np.where(sub == 1, h, np.float32(0))).astype(np.float64)[soil]
    X = xs[soil]
    r2cut = (2.0 * float(p.h)) ** 2 * 1.001
    margin = 0.01 * float(p.h)
    hh = float(p.h)

    def gap(d, fa):
        g = np.where(d > 0, d * hh - fa, np.where(d < 0, fa - (d + 1) * hh, 0.0))
        return np.maximum(g - margin, 0.0)

    tot = np.zeros(len(X), np.int64)
    for dx in range(-2, 3):
        for dy in range(-2, 3):
            cx, cy = X[:, 0] + dx, X[:, 1] + dy
            ok = (cx >= 0) & (cy >= 0) & (cx < 2 * nx) & (cy < 2 * ny)
            rem = r2cut - gap(dx, f[:, 0]) ** 2 - gap(dy, f[:, 1]) ** 2
            ok &= rem > 0
            dz = np.sqrt(np.maximum(rem, 0)) + margin
            zlo = np.maximum(X[:, 2] + np.maximum(-2, np.floor((f[:, 2] - dz) / hh)), 0).astype(np.int64)
            zhi = np.minimum(X[:, 2] + np.minimum(2, np.floor((f[:, 2] + dz) / hh)), ncol2 - 1).astype(np.int64)
            ok &= zlo <= zhi
            base = (np.where(ok, cx, 0) * (2 * ny) + np.where(ok, cy, 0)) * ncol2
            n = pre[np.where(ok, base + zhi + 1, 0)] - pre[np.where(ok, base + zlo, 0)]
            tot += np.where(ok, n, 0)
    return tot


WORKLOADS = {
    "c1": "c1: retaining wall 0.4x0.2x0.3 m backfill, dx=0.005 m, wall translating away at 1 mm/s",
    "c2": "c2: retaining wall 1.0x0.5x0.6 m backfill, dx=0.0063 m, toe rotation 0.01 rad/s (active), "
          "jitter 0.05dx, void ratio U(0.65,0.75)",
    "c2_dx004": "c2 at dx=0.004 m: retaining wall 1.0x0.5x0.6 m backfill, toe rotation 0.01 rad/s (active), "
                "jitter 0.05dx, void ratio U(0.65,0.75)",
    "c3": "c3: retaining wall 2.0x1.0x1.0 m backfill, dx=0.004 m, 15 deg sloping top, 10 kPa surface strip "
          "0.5-1.0 m behind the wall, wall translating away at 1 mm/s",
}


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(case_name, seconds_target=20.0):
    """FP32 oracle on the host cores for a bounded number of whole steps of
    the same case: its per-particle loops run on OpenMP threads (one per
    core the run may use, OMP_NUM_THREADS when set), and one serial step is
    timed beside it."""
    from retainwall_code_amd import cases
    from oracle import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    case = cases.config(case_name)
    p = cases.make_params(case)
    o = O.Oracle(p, np.float32, cases.motion_tables(case)).load_host(case)
    o.abs2rel()
    o.neighbor_search()
    o.sort_arrays_primary()
    print(f"[cpu baseline] {case_name} set up; one serial step", file=sys.stderr, flush=True)
    O.set_threads(1)
    serial = o.run(1)["seconds"]
    print(f"[cpu baseline] serial step {serial:.2f} s; steps on {threads} threads", file=sys.stderr, flush=True)
    O.set_threads(threads)
    r = o.run(1)
    nsteps, secs = 1, r["seconds"]
    if secs < seconds_target / 2:
        more = max(1, int(seconds_target / max(secs, 1e-3)) - 1)
        r = o.run(more)
        nsteps, secs = more, r["seconds"]
    return {"value": case.npm * nsteps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{nsteps} full step(s) of the same {case.npm}-soil-particle case, FP32 oracle, "
                      f"{threads} OpenMP threads",
            "seconds": secs, "serial_value": case.npm / serial}


def timed_steps(ctx, steps, graph):
    """Time `steps` steps with HIP events on the library's stream.  Eager
    runs time every kernel class; graph replays time the interaction only
    (each timed section adds two event-record nodes to the graph, and timing
    every class costs the replay its advantage).  The timed graph is
    captured by two extra untimed steps first, so the timed region holds
    replays only.  Returns (seconds, kernel times)."""
    def arm():
        if graph:
            ctx.enable_timing_classes(["interaction"])
        else:
            ctx.enable_timing(True)
    arm()
    if graph:
        ctx.step(2, graph=True)
    ctx.sync()
    arm()  # resets the counters
    t0 = time.perf_counter()
    ctx.step(steps, graph=graph)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    return elapsed, ctx.kernel_times()


def pick_variant(ctx, log=None):
    """Untimed: the fastest of the bit-identical interaction variants
    (lq_autotune), then the COLUMNS variant (same neighbour sets, other
    summation order) with its own block order and register build, kept if it
    is faster.  Returns (variant, tile_shape, tile_tz, remap)."""
    from retainwall_code_amd import _abi
    variant, tile_shape, tile_tz, remap, ms_exact = ctx.autotune(reps=3)
    best = (variant, tile_shape, tile_tz, remap)
    try:
        ctx.set_interaction(_abi.INTERACTION_COLUMNS, 1)
    except Exception:  # periodic planes, Morton keys or too many cells along z
        return best
    v2, s2, t2, r2, ms_cols = ctx.autotune(reps=3)
    if log:
        log(f"[bench] autotune: variant {variant} {ms_exact:.3f} ms, columns {ms_cols:.3f} ms per step's two calls")
    if ms_cols < ms_exact:
        return v2, s2, t2, r2
    if variant == _abi.INTERACTION_TILE:
        ctx.set_tile(tile_shape, tile_tz)
    ctx.set_interaction(variant, remap)
    return best


def bench_batch(args):
    """N independent cases (sizes log-uniform 5k-200k soil particles, heights,
    materials and wall motions drawn per case, dx = 5 mm) stepped together in
    one batched context; value = total soil particle-steps/s.  For comparison,
    16 of the cases are also stepped alone (one context each) and the time of
    running all N one after another is estimated from a least-squares fit of
    ms/step against particle count."""
    from retainwall_code_amd import _abi, cases, sim
    t0 = time.perf_counter()
    cs = cases.batch_cases(args.batch, seed=4, dr=0.005)
    npm = sum(c.npm for c in cs)
    npb = sum(c.npb for c in cs)
    print(f"[bench] batch of {len(cs)} cases: {npm} soil + {npb} wall particles "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    ctx = sim.open_batch(cs)
    ctx.step(args.warmup, graph=args.graph)
    ctx.sync()
    names = {_abi.INTERACTION_TILE: "tile", _abi.INTERACTION_REFERENCE_ORDER: "reference",
             _abi.INTERACTION_NEIGHBOR_LIST: "neighbor_list", _abi.INTERACTION_PAIR_LIST: "pair_list",
             _abi.INTERACTION_COLUMNS: "columns"}
    if args.variant == "auto":
        variant, shape, tz, remap = pick_variant(ctx, lambda m: print(m, file=sys.stderr, flush=True))
    else:
        variant, remap = {v: k for k, v in names.items()}[args.variant], args.remap
        ctx.set_interaction(variant, remap)
    el, kt = timed_steps(ctx, args.steps, args.graph)
    F = ctx.wall_force_cases([c.probe_box for c in cs])
    ctx.close()
    # sequential one-case contexts on a sample
    rng = np.random.default_rng(1)
    sample = [] if args.no_sample else sorted(rng.choice(len(cs), size=min(16, len(cs)), replace=False))
    xs, ys = [], []
    for k in sample:
        one = sim.open_case(cs[k])
        one.step(2)
        one.sync()
        ta = time.perf_counter()
        one.step(5, graph=args.graph)
        one.sync()
        ys.append(1e3 * (time.perf_counter() - ta) / 5)
        xs.append(cs[k].npm)
        one.close()
    b1, b0 = np.polyfit(np.array(xs, float), np.array(ys, float), 1) if len(xs) > 1 else (float("nan"), float("nan"))
    seq_ms = float(sum(b0 + b1 * c.npm for c in cs))
    out = {
        "metric": "particle-steps/sec", "value": round(npm * args.steps / el, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch of {len(cs)} independent wall cases (5k-200k soil each, dx 5 mm, "
                               "H U(0.1,0.6) m, phi_c U(28,36), e0 U(0.6,0.8), translation or toe rotation)",
                   "cases": len(cs), "soil_particles": npm, "boundary_particles": npb, "graph": bool(args.graph),
                   "interaction_variant": names[variant], "xcd_remap": remap},
        "one_case_per_context": {"sampled_cases": len(sample), "fit_ms_per_step": [round(b0, 4), round(b1, 9)],
                                 "sequential_ms_per_step_estimate": round(seq_ms, 2),
                                 "batch_speedup": round(seq_ms / (1e3 * el / args.steps), 2)},
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1] > 0} if kt else None,
        "wall_force_kN_first_cases": [round(float(f), 6) for f in F[:4]],
    }
    print(json.dumps(out), flush=True)


def surrogate_campaign(rows, epochs, log):
    """Training set of the HypoCS surrogate: exact (x, y) tuples logged from
    runs of other workloads -- c1 (2000 steps), c2 drawn with another seed
    (another jitter and void-ratio realisation, 1000 steps) and c3 (60
    steps) -- rows / 3 from each, spread evenly over each run; then the
    normalisation, the guard box and two members trained on all of them
    (Adam, batch 4096, the library's FP32 training kernels)."""
    from retainwall_code_amd import cases, surrogate
    plan = [("c1", lambda: cases.config("c1"), 2000), ("c2 seed 7", lambda: cases.config("c2", seed=7), 1000),
            ("c3", lambda: cases.config("c3"), 60)]
    xs, ys, info = [], [], []
    for name, make, steps in plan:
        t = time.perf_counter()
        case = make()
        x, y = surrogate.collect(case, steps, rows // len(plan))
        xs.append(x)
        ys.append(y)
        info.append({"source": name, "steps": steps, "soil_particles": case.npm, "tuples": int(len(x)),
                     "seconds": round(time.perf_counter() - t, 1)})
        log(f"[surrogate] logged {len(x)} tuples from {steps} steps of {name} ({info[-1]['seconds']} s)")
        del case
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    del xs, ys
    norm = surrogate.normalisation(x, y)
    t = time.perf_counter()
    members, losses = [], []
    # learning rate 1e-3 for `epochs` epochs, then 3e-4 and 1e-4 for half as many each
    sched = [(epochs, 1e-3), (max(1, epochs // 2), 3e-4), (max(1, epochs // 2), 1e-4)]
    for seed in (1, 2):
        p, hist = None, []
        for k, (ne, lr) in enumerate(sched):
            p, h = surrogate.train(x, y, norm, epochs=ne, batch=4096, lr=lr, seed=seed + 100 * k, params=p)
            hist += [round(float(v), 6) for v in h]
        members.append(p)
        losses.append(hist)
        log(f"[surrogate] member {seed}: loss per epoch {losses[-1]}")
    lo, hi = surrogate.guard_box(x)
    return members, norm, lo, hi, {"sources": info, "tuples": int(len(x)),
                                   "schedule": [{"epochs": ne, "lr": lr} for ne, lr in sched],
                                   "seconds": round(time.perf_counter() - t, 1), "loss_per_epoch": losses}


def _run_chunks(ctx, steps, every, force):
    """Steps in graph-replayed chunks of `every`, wall force after each;
    the first chunk (which captures the graph) is not timed."""
    F, el, timed = [], 0.0, 0
    for k in range(steps // every):
        ctx.sync()
        ta = time.perf_counter()
        ctx.step(every, graph=True)
        ctx.sync()
        if k > 0:
            el += time.perf_counter() - ta
            timed += every
        F.append(force(ctx))
    return np.array(F), el, timed


def bench_surrogate(args):
    """The learned form (config c5): the surrogate trained on other runs
    (surrogate_campaign), then args.sur_steps steps of the case exact and with
    the surrogate (two members, guard box, 1 % audit), wall force every 50
    steps; wall-force error = max_t |F_s - F_p| / max_t |F_p|.  With --batch
    N the two forms run a batch of N cases of the same sand (uneven sizes,
    heights and wall motions) and plain c2 is timed beside them."""
    from retainwall_code_amd import _abi, cases, sim, surrogate

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    members, norm, lo, hi, train = surrogate_campaign(args.sur_rows, args.sur_epochs, log)
    every = 50
    out_extra = {}
    if args.batch:
        cs = cases.batch_cases(args.batch, seed=5, dr=0.005, material=cases.hypo_cs_sand(), void_range=(0.65, 0.75))
        plain, sur = sim.open_batch(cs), sim.open_batch(cs)
        npm = sum(c.npm for c in cs)
        boxes = [c.probe_box for c in cs]
        force = lambda ctx: np.array(ctx.wall_force_cases(boxes))  # noqa: E731
        workload = f"batch of {len(cs)} wall cases of the c2 sand (5k-200k soil each, dx 5 mm), HypoCS surrogate"
        c2 = cases.config("c2")
        ref = sim.open_case(c2)
        ref.step(10, graph=True)
        ref.sync()
        pick_variant(ref, log)
        t = time.perf_counter()
        ref.step(100, graph=True)
        ref.sync()
        out_extra["plain_c2_value"] = round(c2.npm * 100 / (time.perf_counter() - t), 1)
        ref.close()
    else:
        case = cases.config(args.case)
        plain, sur = sim.open_case(case), sim.open_case(case)
        npm = case.npm
        force = lambda ctx: ctx.wall_force(case.probe_box)  # noqa: E731
        workload = f"{args.case} with the HypoCS surrogate (2 members, guard, 1% audit)"
    for ctx in (plain, sur):  # the same interaction variant for both forms
        ctx.step(2, graph=True)
        ctx.sync()
    v = pick_variant(plain, log)
    if v[0] == _abi.INTERACTION_TILE:
        sur.set_tile(v[1], v[2])
    sur.set_interaction(v[0], v[3])
    out_extra["interaction_variant"] = int(v[0])
    surrogate.install(sur, members, norm, lo, hi, ens_tol=args.ens_tol, audit_frac=0.01)
    surrogate.enable(sur)
    surrogate.stats(sur, reset=True)
    res = {}
    for name, ctx in (("exact", plain), ("surrogate", sur)):
        res[name] = _run_chunks(ctx, args.sur_steps, every, force)
        log(f"[surrogate] {name}: {res[name][2]} timed steps in {res[name][1]:.3f} s")
    st = surrogate.stats(sur)
    plain.close()
    sur.close()
    # noise floor of the observable: the same exact steps with another
    # interaction variant (same neighbour sets, other FP32 summation order)
    other = sim.open_batch(cs) if args.batch else sim.open_case(case)
    other.step(2, graph=True)
    alt = _abi.INTERACTION_NEIGHBOR_LIST if v[0] == _abi.INTERACTION_COLUMNS else _abi.INTERACTION_COLUMNS
    try:
        other.set_interaction(alt, 1)
    except Exception:
        alt = _abi.INTERACTION_REFERENCE_ORDER
        other.set_interaction(alt, 1)
    fo = _run_chunks(other, args.sur_steps, every, force)[0]
    other.close()
    log(f"[surrogate] spread run with variant {alt} done")
    fp, fs = res["exact"][0], res["surrogate"][0]

    def rel(a, b):  # max_t |a - b| / max_t |b|, per case when there are several
        return np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-30)

    errs, spread = rel(fs, fp), rel(fo, fp)
    werr = float(np.max(errs))
    out_extra["exact_spread"] = float(np.max(spread))
    out_extra["spread_variant"] = int(alt)
    if fp.ndim == 2:  # per case: error of each case's history, worst case reported
        out_extra["wall_force_error_median_case"] = float(np.median(errs))
        out_extra["exact_spread_median_case"] = float(np.median(spread))
        ok = errs <= np.maximum(3.0 * spread, 0.005)
        out_extra["cases_within_3x_spread_or_0.5pct"] = int(np.count_nonzero(ok))
    else:
        out_extra["wall_force_kN"] = {"exact": [round(float(x), 6) for x in fp],
                                      "surrogate": [round(float(x), 6) for x in fs],
                                      "exact_other_order": [round(float(x), 6) for x in fo]}
    calls = st["surrogate"] + st["fallback_range"] + st["fallback_ensemble"]
    v_exact = npm * res["exact"][2] / res["exact"][1]
    v_sur = npm * res["surrogate"][2] / res["surrogate"][1]
    out = {
        "metric": "particle-steps/sec", "value": round(v_sur, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": res["surrogate"][2], "warmup": every, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (surrogate bf16 MFMA)", "data": "synthetic",
        "config": {"workload": workload, "soil_particles": npm, "graph": True, "force_every": every},
        "exact_value": round(v_exact, 1), "gain": round(v_sur / v_exact, 3),
        "wall_force_error": werr, "accepted_fraction": round(st["surrogate"] / max(calls, 1), 4), "stats": st,
        "training": train,
    }
    if "plain_c2_value" in out_extra:
        out["gain_vs_plain_c2"] = round(v_sur / out_extra["plain_c2_value"], 3)
    out.update(out_extra)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--eager", action="store_true",
                    help="enqueue every kernel per step instead of replaying the captured two-step hipGraph")
    ap.add_argument("--variant", default="auto",
                    choices=["auto", "tile", "reference", "neighbor_list", "pair_list", "columns"],
                    help="interaction kernel variant (columns sums each particle's pairs in another order; the others are bit-identical); auto times each and keeps the fastest")
    ap.add_argument("--remap", type=int, default=0, help="XCD block renumbering for reference/neighbor_list")
    ap.add_argument("--key", default="linear", choices=["linear", "morton"], help="cell-key function")
    ap.add_argument("--tile", default="", metavar="SHAPE,TZ", help="fix the tile shape/depth instead of autotuning")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="batched workload: N independent wall cases in one context (256 = the c4 workload)")
    ap.add_argument("--no-sample", action="store_true",
                    help="with --batch: skip the one-case-per-context comparison runs")
    ap.add_argument("--surrogate", action="store_true",
                    help="gain test of the HypoCS surrogate on the case: log, train, then time both forms")
    ap.add_argument("--ens-tol", type=float, default=0.1, help="surrogate ensemble tolerance (normalised units)")
    ap.add_argument("--sur-epochs", type=int, default=4, help="surrogate training epochs per member")
    ap.add_argument("--sur-rows", type=int, default=50_000_000, help="surrogate training tuples (logged from c1-c3)")
    ap.add_argument("--sur-steps", type=int, default=1000, help="surrogate comparison steps (force every 50)")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="BYTES",
                    help="only run the STREAM copy probe over BYTES (HBM bandwidth / counter calibration)")
    ap.add_argument("--probe-mlp", type=int, default=0, metavar="ROWS",
                    help="only run the BF16 surrogate forward on ROWS random rows, 6 times (k_mlp_bf16 timing)")
    args = ap.parse_args()
    args.graph = not args.eager
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU: independent cases shard across "
                                   "GPUs with no exchange and a single case would shard by y-slabs with a 2h "
                                   "halo, but multi-GPU runs are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from retainwall_code_amd import cases, sim

    if args.probe_copy:
        ctx = sim.open_case(cases.config("tiny"))
        ms = ctx.stream_copy(args.probe_copy, reps=20)
        print(json.dumps({"probe": "stream_copy", "bytes_read": args.probe_copy, "bytes_written": args.probe_copy,
                          "ms_per_launch": ms, "GBps": 2 * args.probe_copy / (ms * 1e-3) / 1e9}), flush=True)
        ctx.close()
        return

    if args.probe_mlp:
        from retainwall_code_amd import surrogate
        rng = np.random.default_rng(0)
        p = (rng.standard_normal(surrogate.NPARAM) * 0.1).astype(np.float32)
        norm = np.concatenate([np.zeros(surrogate.NIN), np.ones(surrogate.NIN), np.zeros(surrogate.NOUT),
                               np.ones(surrogate.NOUT)]).astype(np.float32)
        x = rng.standard_normal((args.probe_mlp, surrogate.NIN)).astype(np.float32)
        t = time.perf_counter()
        for _ in range(6):
            y = surrogate.forward(p, norm, x, bf16=True)
        print(json.dumps({"probe": "mlp_bf16_forward", "rows": args.probe_mlp, "calls": 6,
                          "flops_per_row": 2 * (surrogate.NIN * surrogate.NH + 2 * surrogate.NH * surrogate.NH +
                                                surrogate.NH * surrogate.NOUT),
                          "host_seconds_incl_copies": round(time.perf_counter() - t, 3),
                          "mean_abs_y": float(np.abs(y).mean())}), flush=True)
        return
    if args.surrogate:
        return bench_surrogate(args)
    if args.batch:
        return bench_batch(args)

    case = cases.config(args.case)
    print(f"[bench] case={args.case} np={case.np} npb={case.npb} npm={case.npm}", file=sys.stderr, flush=True)
    from retainwall_code_amd import _abi
    names = {_abi.INTERACTION_TILE: "tile", _abi.INTERACTION_REFERENCE_ORDER: "reference",
             _abi.INTERACTION_NEIGHBOR_LIST: "neighbor_list", _abi.INTERACTION_PAIR_LIST: "pair_list",
             _abi.INTERACTION_COLUMNS: "columns"}
    ids = {v: k for k, v in names.items()}
    variant = ids.get(args.variant, _abi.INTERACTION_REFERENCE_ORDER)
    key_mode = _abi.KEY_MORTON if args.key == "morton" else _abi.KEY_LINEAR
    ctx = sim.open_case(case, key_mode=key_mode, variant=variant)
    p = ctx.params
    ctx.step(args.warmup, graph=args.graph)
    ctx.sync()
    tile_shape = tile_tz = None
    remap = args.remap
    if args.variant == "tile" and args.tile:
        tile_shape, tile_tz = (int(v) for v in args.tile.split(","))
        ctx.set_tile(tile_shape, tile_tz)
    elif args.variant == "auto":
        variant, tile_shape, tile_tz, remap = pick_variant(ctx, lambda m: print(m, file=sys.stderr, flush=True))
    elif args.variant == "columns":  # untimed: block order and register build of the COLUMNS kernel
        ctx.set_interaction(variant, remap)
        variant, tile_shape, tile_tz, remap, _ = ctx.autotune(reps=3)
    else:
        ctx.set_interaction(variant, remap)
    if variant != _abi.INTERACTION_TILE:
        tile_shape = tile_tz = None
    vname = names[variant]
    # interaction work per call (outside the timed region)
    soil = slice(case.npb, case.np)
    cellpos = ctx.get("cellpos")
    beg, end = ctx.get("cell_beg"), ctx.get("cell_end")
    cand = candidate_counts(cellpos, beg, end, p, soil)
    if variant == _abi.INTERACTION_COLUMNS:
        cand = col_candidate_counts(ctx.get("pos"), cellpos, p, soil, case.np)
    nbc, _ = ctx.neighbor_lists(1)
    flops_call = FLOP_CAND * float(cand.sum()) + FLOP_PAIR * float(nbc.sum())
    ncell_linear = 2 * int(p.cell_num[0]) * int(p.cell_num[1]) * int(p.cell_num[2]) + 1
    bytes_call = B_SOIL * case.npm + B_BND * case.npb + B_CELL * ncell_linear

    elapsed, kt = timed_steps(ctx, args.steps, args.graph)
    F = ctx.wall_force(case.probe_box)
    ms_step = 1e3 * elapsed / args.steps
    value = case.npm * args.steps / elapsed

    roof = None
    if kt and kt["interaction"][1] > 0:
        avg_ms = kt["interaction"][0] / kt["interaction"][1]
        achieved = bytes_call / (avg_ms * 1e-3) / 1e9
        traffic = None
        pmc = os.path.join(ROOT, "profiles", f"interaction_pmc_{args.case}.json")
        if os.path.exists(pmc):
            rec = json.load(open(pmc))
            if rec.get("kernel") == {"tile": "k_interaction_tile", "reference": "k_interaction_ref",
                                     "neighbor_list": "k_interaction_nl",
                                     "pair_list": "k_interaction_plist", "columns": "k_interaction_col"}[vname] \
                    and rec.get("key_mode", "linear") == args.key:
                traffic = rec.get("hbm_bytes_per_launch")  # rocprofv3 FETCH/WRITE_SIZE, copy-calibrated
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": {"tile": "k_interaction_tile", "reference": "k_interaction_ref",
                           "neighbor_list": "k_interaction_nl", "pair_list": "k_interaction_plist",
                           "columns": "k_interaction_col"}[vname],
                "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": bytes_call,
                "valu_tflops": round(flops_call / (avg_ms * 1e-3) / 1e12, 2),
                "valu_frac": round(flops_call / (avg_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
                "flops_per_launch": flops_call}
        if traffic:  # measured HBM bytes per launch at the measured launch time
            roof["traffic_GBps"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
            roof["traffic_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.case)
    out = {
        "metric": "particle-steps/sec", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS.get(args.case, f"{args.case} (tests' case generator)"),
                   "soil_particles": case.npm, "boundary_particles": case.npb, "cells": int(p.ncell),
                   "dt": float(p.dt) if p.dt else case.params["dt"], "graph": bool(args.graph),
                   "interaction_variant": vname, "tile_shape": tile_shape, "tile_tz": tile_tz,
                   "xcd_remap": remap if variant != _abi.INTERACTION_TILE else None,
                   "key_mode": args.key},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1] > 0} if kt else None,
        "wall_force_kN": F,
        "mean_neighbors": round(float(nbc.mean()), 2), "mean_candidates": round(float(cand.mean()), 2),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
