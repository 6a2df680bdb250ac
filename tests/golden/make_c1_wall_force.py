"""Write the c1 wall earth-pressure histories of the CPU oracle.

Workload: the small retaining-wall case (0.4 x 0.2 x 0.3 m backfill at
dx = 5 mm, 250,240 soil particles on the GenCase lattice, rigid vertical wall
translating away from the soil at 1 mm/s), 2000 steps, the wall force
(cuGetValuesAtPoint, utilities.cu:578-670) sampled every 50 steps as the
reference's time loop does at its saves (loquat.cu:548-561).

Two lattices: the exact GenCase lattice, and the same lattice jittered by
U(-0.05 dx, 0.05 dx) per axis (seed 2024).  Each is run by the FP64 oracle
(the accuracy reference) and by the FP32 oracle (the reference's own FP32
operation order).  The oracle's per-particle loops run on OpenMP threads;
its results do not depend on the thread count.

    python tests/golden/make_c1_wall_force.py            # all four runs
    python tests/golden/make_c1_wall_force.py --check    # first sample only

Output: tests/golden/c1_wall_force.json.  Run time on 8 cores: about
10 minutes per run.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "c1_wall_force.json")
STEPS = 2000
EVERY = 50
LATTICES = {"exact": {}, "jitter": {"jitter": 0.05, "seed": 2024}}


def case_for(lattice):
    from retainwall_code_amd import cases
    return cases.config("c1", **LATTICES[lattice])


def case_digest(case) -> str:
    """sha256 over the generated initial state (positions, velocities,
    masses, codes, stresses): the fixture describes exactly this case."""
    h = hashlib.sha256()
    for a in (case.pospres, case.velrhop, case.idmassidx, case.codes, case.str):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def oracle_for(case, dtype):
    """Oracle loaded with the case and the initial search and sort
    (loquat.cu:528-531), as the GPU context starts."""
    from oracle import oracle as O
    from retainwall_code_amd import cases
    o = O.Oracle(cases.make_params(case), dtype, cases.motion_tables(case)).load_host(case)
    o.abs2rel()
    o.neighbor_search()
    o.sort_arrays_primary()
    return o


def run(lattice, dtype, steps=STEPS, every=EVERY):
    case = case_for(lattice)
    o = oracle_for(case, dtype)
    t0 = time.time()
    r = o.run(steps, case.probe_box, every)
    return case, r["forces"], time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true", help="recompute only the first sample of each history")
    a = ap.parse_args()
    if a.check:
        d = json.load(open(OUT))
        for lat in LATTICES:
            _, f, s = run(lat, np.float64, EVERY, EVERY)
            want = d["lattices"][lat]["f64"][0]
            print(f"{lat}: first FP64 sample {f[0]!r} (fixture {want!r}) in {s:.1f} s")
        return
    out = {"workload": "c1: 0.4x0.2x0.3 m backfill, dx 0.005 m, wall translating away at 1 mm/s",
           "steps": STEPS, "every": EVERY, "unit": "kN (0.001 * sum of contact_mag.xx, utilities.cu:592-667)",
           "lattices": {}}
    for lat in LATTICES:
        row = {}
        for name, dt in (("f64", np.float64), ("f32", np.float32)):
            case, f, s = run(lat, dt)
            print(f"{lat} {name}: {len(f)} samples in {s:.0f} s, last {f[-1]:.6e}", flush=True)
            row[name] = [float(x) for x in f]
        row.update({"np": int(case.np), "npb": int(case.npb), "case_sha256": case_digest(case),
                    "params": LATTICES[lat], "probe_box": [float(x) for x in case.probe_box]})
        out["lattices"][lat] = row
        with open(OUT, "w") as fh:
            json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
