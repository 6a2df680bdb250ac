"""Reference case files: <name>.xml, <name>.dat, <name>.motion.

Writers follow GenCase.cu:70-227 (SaveDatFile, SaveXmlFile, SaveMotionFile):
same element names and type attributes, same column order (boundary rows
with a 4-column normal, then soil rows), same motion-table layout.  The
reader mirrors LoquatIo (loquat_io.cu:75-312) for tests.
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import cases


def _g(v):
    """C++ ostream default formatting (6 significant digits)."""
    return f"{v:g}" if isinstance(v, float) or isinstance(v, np.floating) else str(v)


def write_case(case: cases.Case, directory: str, name: str, time_max=None, output_frequency=None,
               output_format=1) -> str:
    os.makedirs(directory, exist_ok=True)
    base = os.path.join(directory, name)
    p = case.params
    np_ = case.np
    with open(base + ".dat", "w") as f:
        for i in range(np_):
            c = case.codes[i]
            row = [str(int(x)) for x in c]
            row += [repr(float(x)) for x in case.pospres[i]]
            row += [repr(float(x)) for x in case.velrhop[i]]
            row += [repr(float(case.idmassidx[i, 0])), repr(float(case.idmassidx[i, 1])),
                    repr(float(case.idmassidx[i, 3]))]
            row += [repr(float(x)) for x in case.str[i]]
            row += [repr(float(x)) for x in case.stn[i]]
            if i < case.npb:
                row += [repr(float(x)) for x in case.normal[i]]
            f.write(" ".join(row) + "\n")
    tm = p["time_max"] if time_max is None else time_max
    of = p["output_frequency"] if output_frequency is None else output_frequency
    items = [("dimension-of-problem", "dimension", p["dim"]), ("number-of-motion", "int", p["n_motion"]),
             ("period-bound", "bool", p["periodbound"])]
    if p["periodbound"]:  # GenCase.cu:134-143
        for side in ("left", "right"):
            for k, ab in enumerate("ABCD"):
                items.append((f"plane{side}-{ab}", "scalar", repr(float(p["plane" + side][k]))))
    items += [("number-of-particles", "int", np_),
              ("number-of-boundary-particles", "int", case.npb), ("number-of-material-particles", "int", case.npm)]
    for k, ax in enumerate("xyz"):
        items.append((f"domain-size-{ax}", "scalar", repr(float(p["domain_size"][k]))))
    for k, ax in enumerate("xyz"):
        items.append((f"domain-min-{ax}", "scalar", repr(float(p["domain_min"][k]))))
    items += [("particle-size", "scalar", repr(float(p["dr"]))), ("h-over-dr", "scalar", repr(float(p["hdr"])))]
    for k, ax in enumerate("xyz"):
        items.append((f"acceleration-{ax}", "scalar", repr(float(p["gravity"][k]))))
    items += [("reference-density-material", "scalar", repr(float(p["rho0_m"]))),
              ("damping", "scalar", repr(float(p["damping"]))), ("number-of-model", "int", p["n_model"])]
    for i in range(p["n_model"]):
        for j in range(15):
            items.append((f"model-{i}-par-{j}", "scalar", repr(float(p["model_par"][i][j]))))
    items += [("stress-shepard-filter", "int", p["stress_shepard_filter"]),
              ("dynamic-time-step", "int", p["dynamic_time_step"]),
              ("integration-method", "int", p["integration_method"]),
              ("viscosity-type", "int", p["viscosity_type"])]
    if p["viscosity_type"] == 1:
        items.append(("artificial-viscosity-alpha-material", "scalar", repr(float(p["artificial_viscosity_alpha_m"]))))
    items += [("speed-of-sound-material", "scalar", repr(float(p["cs_m"]))),
              ("cfl-number", "scalar", repr(float(p["cfl"]))),
              ("artificial-pressure-coefficient-material", "scalar",
               repr(float(p["artificial_pressure_coefficient_material"]))),
              ("time-max", "scalar", repr(float(tm))), ("time", "scalar", repr(float(p["time"]))),
              ("time-step", "scalar", repr(float(p["dt"]))), ("output-frequency", "scalar", repr(float(of))),
              ("output-number", "int", p["output_number"]), ("output-format", "int", output_format)]
    with open(base + ".xml", "w") as f:
        f.write('<?xml version="1.0" encoding="utf-8" ?>\n<parameters>\n')
        for k, t, v in items:
            f.write(f'  <{k} type="{t}">{v}</{k}>\n')
        f.write("</parameters>\n")
    if p["n_motion"] > 0:
        with open(base + ".motion", "w") as f:
            for t in case.motion:
                f.write(f"{t['is_translation']} {t['is_rotation']}\n")
                if t["is_translation"]:
                    f.write(f"{t['table_size_trans']}\n")
                    for j in range(t["table_size_trans"]):
                        v, a = t["trans_vel"][j], t["trans_acc"][j]
                        f.write(" ".join(repr(float(x)) for x in [t["trans_time"][j], *v, *a]) + "\n")
                if t["is_rotation"]:
                    f.write(" ".join([str(t["pattern"])] + [repr(float(x)) for x in [*t["pos_rp"], *t["rotat_vel"]]]
                                     + [str(t["table_size_rotat"])]) + "\n")
                    for j in range(t["table_size_rotat"]):
                        f.write(" ".join(repr(float(x)) for x in [t["rotat_time"][j], *t["rotat_acc"][j]]) + "\n")
    return base


def read_xml(path: str) -> dict:
    txt = open(path).read()
    out = {}
    for m in re.finditer(r"<([A-Za-z0-9_\-]+)\s+type=\"([a-z_]+)\">([^<]*)</\1>", txt):
        k, t, v = m.groups()
        out[k] = int(v) if t in ("int", "dimension", "model_mat") else (bool(float(v)) if t == "bool" else float(v))
    return out


def read_dat(path: str, npb: int, npm: int) -> dict:
    rows = open(path).read().split("\n")
    b = np.array([r.split() for r in rows[:npb]], np.float64) if npb else np.zeros((0, 31))
    s = np.array([r.split() for r in rows[npb:npb + npm]], np.float64) if npm else np.zeros((0, 27))
    allr = np.concatenate([b[:, :27], s], 0) if npb else s
    return {"codes": allr[:, 0:4].astype(np.int64), "pospres": allr[:, 4:8], "velrhop": allr[:, 8:12],
            "idm": allr[:, 12:15], "str": allr[:, 15:21], "stn": allr[:, 21:27],
            "normal": b[:, 27:31] if npb else np.zeros((0, 4))}
