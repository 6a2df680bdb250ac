"""Retaining-wall case generator (GenCase case 1 semantics, parameterised).

Follows GenCase.cu:229-259 (cell-centred lattice, ``int(L/dr)`` points per
axis, float32 arithmetic as the reference's ``scalar``) and GenCase.cu:352-534
(bottom / bed / moving wall 10 / fixed contact wall 11, soil block at the full
domain width in y, K0 geostatic stress, HypoCS material pack).  Extensions
(with no reference code): jittered positions, per-particle void ratio, an
explicit translation or toe-rotation table, and seeded randomness from a
counter-based splitmix64 so that any implementation can regenerate a case.

A case is a plain dict of numpy arrays plus an ``lq_params``-shaped dict; the
writers/readers for the reference file formats live in ``caseio``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi

f32 = np.float32

# HypoCS parameter map, GenCase.cu:52-68.
HYPO_KEYS = ["Ei", "Vi", "phi_c", "cohesion", "p_ref", "void_co", "lambda", "xi", "alpha",
             "void_min", "void_max", "void_0", "rho0", "rho_min", "rho_max"]


def hypo_cs_sand(rho0=1600.0, phi_c_deg=30.0, e0=0.70, h_s=5.8e9, n=0.28, e_d0=0.53, e_c0=0.84,
                 e_i0=1.0, alpha=0.13, Ei=5.0e7, Vi=0.3, p_ref=1.0e5):
    """von Wolffersdorff-style sand parameters mapped onto HypoCS.

    e_c = e_c0 exp(-(3p/h_s)^n) equals the model's e_c0 exp(-lambda (p/p_ref)^xi)
    (constitutive_model.h:262) for xi = n and lambda = (3 p_ref / h_s)^n.  e_d0
    becomes the constant minimum void ratio and e_i0 the clamp maximum; beta has
    no counterpart.  Ei, Vi follow GenCase.cu:387.
    """
    lam = (3.0 * p_ref / h_s) ** n
    return [Ei, Vi, phi_c_deg * math.pi / 180.0, 0.0, p_ref, e_c0, lam, n, alpha, e_d0, e_i0, e0,
            rho0, rho0 * 0.8, rho0 * 1.2]


def dp_soil(E=2.0e7, V=0.3, phi_deg=30.0, psi_deg=0.0, cohesion=0.0, rho0=1600.0):
    """Drucker-Prager row, GetMaterial_DP_EPP (GenCase.cu:16-25): bulk and
    shear modulus, then k_phi, k_psi and k_c of the cone fitted to
    Mohr-Coulomb (plane strain); angles in degrees here."""
    tphi, tpsi = math.tan(math.radians(phi_deg)), math.tan(math.radians(psi_deg))
    row = [0.0] * 15
    row[0] = E / (3.0 * (1.0 - 2.0 * V))
    row[1] = E / (2.0 * (1.0 + V))
    row[2] = 3.0 * tphi / math.sqrt(9.0 + 12.0 * tphi * tphi)
    row[3] = 3.0 * tpsi / math.sqrt(9.0 + 12.0 * tpsi * tpsi)
    row[4] = 3.0 * cohesion / math.sqrt(9.0 + 12.0 * tphi * tphi)
    row[13], row[14] = rho0 * 0.8, rho0 * 1.2
    return row


def mui_dp_soil(E=2.0e7, V=0.3, phi1_deg=30.0, phi2_deg=40.0, I0=0.3, rho_g=2650.0, d_grain=0.001,
                lam=1.0e-3, rho0=1600.0):
    """mu(I) Drucker-Prager row, GetMaterial_MuI_DP_EPP (GenCase.cu:27-40):
    moduli, mu_s = tan(phi1), mu_2 = tan(phi2), I0, grain density, grain
    diameter and the regularisation lambda."""
    row = [0.0] * 15
    row[0] = E / (3.0 * (1.0 - 2.0 * V))
    row[1] = E / (2.0 * (1.0 + V))
    row[2] = math.tan(math.radians(phi1_deg))
    row[3] = math.tan(math.radians(phi2_deg))
    row[4], row[5], row[6], row[7] = I0, rho_g, d_grain, lam
    row[13], row[14] = rho0 * 0.8, rho0 * 1.2
    return row


def hypo_basic_soil(Ei=5.0e7, Vi=0.3, phi_c_deg=30.0, cohesion=0.0, p_ref=1.0e5, rho0=1600.0):
    """Basic hypoplastic row, GetMaterial_HypoBasic (GenCase.cu:42-50)."""
    row = [0.0] * 15
    row[0], row[1], row[2], row[3], row[4] = Ei, Vi, math.radians(phi_c_deg), cohesion, p_ref
    row[13], row[14] = rho0 * 0.8, rho0 * 1.2
    return row


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, ids: np.ndarray, stream: int) -> np.ndarray:
    """Counter-based U[0,1): splitmix64(seed, stream, id) -> top 53 bits."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream) * np.uint64(0x9E3779B97F4A7C15)
        z = _splitmix64(ids.astype(np.uint64) ^ base)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Case:
    params: dict
    codes: np.ndarray
    pospres: np.ndarray
    velrhop: np.ndarray
    idmassidx: np.ndarray
    str: np.ndarray
    stn: np.ndarray
    normal: np.ndarray
    motion: list = field(default_factory=list)
    probe_box: tuple = (0.0, 0.0, 0.0)
    name: str = "case"
    surcharge: tuple = None  # (soil block id, pressure Pa) of a strip load, or None

    @property
    def np(self):
        return self.params["np"]

    @property
    def npb(self):
        return self.params["npb"]

    @property
    def npm(self):
        return self.params["npm"]


def motion_table(mode: str, speed: float, pos_rp=(0.0, 0.0, 0.0), t_end=1.0e6):
    """One rigid-body table: 'translate' (vx = speed) or 'rotate' (wy = speed)."""
    t = {"is_translation": 0, "table_size_trans": 0, "trans_vel": [[0.0] * 3] * 10,
         "trans_time": [0.0] * 10, "trans_acc": [[0.0] * 3] * 10, "is_rotation": 0,
         "pattern": _abi.CUBOID, "table_size_rotat": 0, "pos_rp": list(pos_rp),
         "rotat_vel": [0.0] * 3, "rotat_time": [0.0] * 10, "rotat_acc": [[0.0] * 3] * 10}
    t = {k: (list(map(list, v)) if isinstance(v, list) and v and isinstance(v[0], list) else
             (list(v) if isinstance(v, list) else v)) for k, v in t.items()}
    if mode == "translate":
        t["is_translation"] = 1
        t["table_size_trans"] = 1
        t["trans_time"][0] = t_end
        t["trans_vel"][0] = [speed, 0.0, 0.0]
    elif mode == "rotate":
        t["is_rotation"] = 1
        t["table_size_rotat"] = 1
        t["rotat_time"][0] = t_end
        t["rotat_vel"] = [0.0, speed, 0.0]
    else:
        raise ValueError(mode)
    return t


def retaining_wall(box_length=0.4, box_width=0.2, box_height=0.3, dr=0.005, hdr=1.3, *,
                   fix_wall_height=0.04, wall_extra=0.05, rho0=1600.0, material=None,
                   cs=65.0, cfl=0.1, alpha_visc=1.5, art_p=0.01, gravity_z=-9.81, K0=0.5,
                   motion=("translate", -0.001), jitter=0.0, void_range=None, seed=0,
                   time_max=1.0, shepard=0, model_type=_abi.HYPO_CS, periodic_y=False, slope_deg=0.0,
                   surcharge=None, name="retaining_wall") -> Case:
    """Build a GenCase case-1 style wall case.

    box_* are the backfill dimensions (x away from the wall, y along it, z up).
    The moving wall (block 10) is box_height + wall_extra tall (GenCase uses
    1.05 for a 1.0 box).  jitter is the half-width of a uniform position
    perturbation in units of dr (soil and walls alike).  void_range=(lo, hi)
    draws each soil particle's initial void ratio e_i and sets its density to
    rho0 (1 + e0) / (1 + e_i) (constitutive_model.h:259 read backwards).
    model_type is the soil's constitutive model (codes.z, types.h:51); pass
    the matching parameter row as material (dp_soil, mui_dp_soil,
    hypo_basic_soil) for any model other than the default HypoCS.
    periodic_y makes the case periodic along y (the wall-parallel direction,
    which has no walls): planes y = 0 and y = npy dr, one lattice period.
    slope_deg tilts the backfill top, rising away from the wall, about the
    middle of the backfill (the mean height stays box_height).  surcharge =
    (pressure, x0, x1) puts a strip load on the free surface between x0 and
    x1 behind the wall face: the top soil particle of each lattice column
    there gets block id 7 and carries pressure * dr^2 (lq_set_surcharge), and
    the initial stress below it includes the load (K0 laterally).
    """
    dr = f32(dr)
    hdr = f32(hdr)
    bt = f32(math.ceil(float(f32(2.0) * hdr))) * dr  # GenCase.cu:283
    vol = dr * dr * dr
    fwh = f32(fix_wall_height)
    mwh = f32(box_height + wall_extra)
    bl, bw, bh = f32(box_length), f32(box_width), f32(box_height)
    dom_x = bl + fwh + f32(2.0) * bt  # GenCase.cu:366
    dom_y = bw + f32(2.0) * bt
    tan_s = math.tan(math.radians(slope_deg))
    soil_top = float(bh) + tan_s * float(bl) / 2.0  # highest backfill point
    dom_z = f32(float(max(float(mwh), soil_top) + fwh) + 2.0 * float(bt))  # 2.0 is a double literal there
    npx = max(int(dom_x / dr), 1)
    npy = max(int(dom_y / dr), 1)
    npz = int(dom_z / dr)
    half = f32(0.5) * dr
    ii, jj, kk = np.meshgrid(np.arange(npx, dtype=np.float32), np.arange(npy, dtype=np.float32),
                             np.arange(npz, dtype=np.float32), indexing="ij")
    x = (half + dr * ii.ravel()).astype(np.float32)
    y = (half + dr * jj.ravel()).astype(np.float32)
    z = (half + dr * kk.ravel()).astype(np.float32)
    del ii, jj, kk
    two = f32(2.0)
    bottom = (z <= bt) & (x >= fwh)
    bed = ((z > fwh) & (z <= fwh + bt) & (x >= 0) & (x <= fwh)) | \
          ((z > bt) & (z <= fwh + bt) & (x >= fwh) & (x <= fwh + bt))
    left = (z > fwh + bt) & (z <= fwh + bt + mwh) & (x >= fwh) & (x <= fwh + bt)
    right = (z >= bt) & (z <= fwh + bt + mwh) & (x >= fwh + bl + bt) & (x <= fwh + bl + two * bt)
    bmask = bottom | bed | left | right
    if slope_deg:
        ztop = (float(bh + fwh + bt) + tan_s * ((x - float(fwh + bt)) - float(bl) / 2.0)).astype(np.float32)
        specimen = (z > bt) & (z <= ztop) & (x > fwh + bt) & (x <= fwh + bt + bl)
    else:
        specimen = (z > bt) & (z <= bh + fwh + bt) & (x > fwh + bt) & (x <= fwh + bt + bl)
    bidx = np.nonzero(bmask)[0]
    sidx = np.nonzero(specimen)[0]
    npb, npm = len(bidx), len(sidx)
    n = npb + npm
    codes = np.zeros((n, 4), np.int32)
    normal = np.zeros((npb, 4), np.float32)
    blk = np.where(bottom[bidx], 0, np.where(bed[bidx], 1, np.where(left[bidx], 10, 11)))
    codes[:npb, 0] = blk
    codes[:npb, 1] = np.where(blk == 10, _abi.NONFIX, _abi.FIX)
    codes[:npb, 2] = np.iinfo(np.int32).min
    codes[:npb, 3] = np.where(blk == 10, 0, np.iinfo(np.int32).min)
    normal[blk == 0] = (0, 0, 1, -1)
    normal[blk == 1] = (1, 0, 0, -1)
    normal[blk == 10] = (1, 0, 0, 1)
    normal[blk == 11] = (-1, 0, 0, 1)
    codes[npb:] = (6, _abi.NONFIX, int(model_type), 0)
    pos = np.zeros((n, 4), np.float32)
    pos[:npb, 0], pos[:npb, 1], pos[:npb, 2] = x[bidx], y[bidx], z[bidx]
    pos[npb:, 0], pos[npb:, 1], pos[npb:, 2] = x[sidx], y[sidx], z[sidx]
    rho0 = f32(rho0)
    # K0 geostatic stress from the untouched lattice height (GenCase.cu:483-485)
    zs = z[sidx]
    if slope_deg:
        szz = (rho0 * f32(gravity_z) * (-zs + ztop[sidx])).astype(np.float32)
    else:
        szz = (rho0 * f32(gravity_z) * (-zs + bt + bh + fwh)).astype(np.float32)
    sur = None
    if surcharge is not None:
        q, sx0, sx1 = surcharge
        xs = x[sidx] - float(fwh + bt)
        strip = (xs >= sx0) & (xs <= sx1)
        col = (np.round((x[sidx] - half) / dr).astype(np.int64) * npy
               + np.round((y[sidx] - half) / dr).astype(np.int64))
        top = np.full(col.max() + 1, -np.inf, np.float32)
        np.maximum.at(top, col, zs)
        surface = strip & (zs == top[col])
        codes[npb:, 0][surface] = 7
        szz = np.where(strip, szz - f32(q), szz).astype(np.float32)
        sur = (7, float(q))
    stress = np.zeros((n, 6), np.float32)
    stress[npb:, 0] = szz * f32(K0)
    stress[npb:, 1] = szz * f32(K0)
    stress[npb:, 2] = szz
    vel = np.zeros((n, 4), np.float32)
    idm = np.zeros((n, 4), np.float32)
    mat = list(material) if material is not None else hypo_cs_sand(rho0=float(rho0))
    rho = np.full(npm, rho0, np.float32)
    if void_range is not None:
        e0 = mat[11]
        u = uniform01(seed, np.arange(npm), 7)
        e_i = void_range[0] + (void_range[1] - void_range[0]) * u
        rho = (float(rho0) * (1.0 + e0) / (1.0 + e_i)).astype(np.float32)
    vel[npb:, 3] = rho
    idm[npb:] = np.stack([np.full(npm, 100, np.float32), (rho * vol).astype(np.float32),
                          np.zeros(npm, np.float32), np.full(npm, 100, np.float32)], 1)
    if jitter:
        ids = np.arange(n)
        for ax in range(3):
            u = uniform01(seed, ids, ax)
            pos[:, ax] += ((2.0 * u - 1.0) * jitter * float(dr)).astype(np.float32)
    mode, speed = motion
    rp = (float(fwh + bt), 0.0, float(fwh + bt))  # toe of the moving wall, GenCase.cu:524
    mt = motion_table(mode, speed, rp)
    dt = float(f32(cfl) * (f32(dr * hdr) / f32(cs)))
    params = {
        "dim": 3, "periodbound": int(bool(periodic_y)), "n_motion": 1, "np": n, "npb": npb, "npm": npm,
        "np_delete_m": 0, "domain_size": [float(dom_x), float(dom_y), float(dom_z)],
        "domain_min": [0.0, 0.0, 0.0], "dr": float(dr), "hdr": float(hdr),
        "gravity": [0.0, 0.0, gravity_z], "rho0_m": float(rho0), "damping": 0.0, "n_model": 1,
        "model_par": [mat] + [[0.0] * 15] * 9, "stress_shepard_filter": shepard,
        "dynamic_time_step": 1, "integration_method": 1, "viscosity_type": 1,
        "artificial_viscosity_alpha_m": alpha_visc, "cs_m": cs, "cfl": cfl,
        "artificial_pressure_coefficient_material": art_p, "time_max": time_max, "time": 0.0,
        "dt": dt, "output_frequency": time_max / 100.0, "output_number": 0, "output_format": 2,
    }
    if periodic_y:  # planeleft n = +y through y = 0, planeright n = -y through y = npy dr
        params["planeleft"] = [0.0, 1.0, 0.0, 0.0]
        params["planeright"] = [0.0, -1.0, 0.0, float(f32(npy * float(dr)))]
    # wall-force probe box (utilities.cu:619-621 with the case's own numbers)
    box = (float(fwh + bt + f32(3.0) * dr), float(bt + fwh), float(bt + fwh + f32(1.0)))
    return Case(params, codes, pos, vel, idm, stress, np.zeros((n, 6), np.float32), normal,
                [mt], box, name, sur)


def make_params(case: Case, key_mode=_abi.KEY_LINEAR, variant=_abi.INTERACTION_NEIGHBOR_LIST,
                lib=None) -> _abi.Params:
    """Fill an lq_params struct from a case and initialise derived fields."""
    p = _abi.Params()
    d = case.params
    for k, v in d.items():
        if k == "model_par":
            for i in range(len(v)):
                for j in range(15):
                    p.model_par[i][j] = v[i][j]
        elif isinstance(v, (list, tuple)):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    for k, t in enumerate(case.motion):
        m = p.mv[k]
        m.is_trans = t["is_translation"]
        m.is_rotat = t["is_rotation"]
        if t["is_translation"]:
            m.trans_vel[:] = t["trans_vel"][0]
        if t["is_rotation"]:
            m.pattern = t["pattern"]
            m.pos_rp[:] = t["pos_rp"]
            m.rotat_vel[:] = t["rotat_vel"]
    p.key_mode = key_mode
    p.interaction_variant = variant
    if lib is None:
        initialize_parameters_py(p)
    else:
        st = lib.lq_initialize_parameters(p)
        if st:
            raise RuntimeError(f"lq_initialize_parameters failed: {st}")
    return p


def initialize_parameters_py(p):
    """loquat.cu:130-190 for the 3D case (float arithmetic as the reference)."""
    p.h = f32(p.dr) * f32(p.hdr)
    p.i_h = f32(1.0) / f32(p.h)
    p.eta = f32(0.01) * f32(p.h) * f32(p.h)
    p.bound_thickness = f32(math.ceil(2.0 * float(p.hdr)) * float(p.dr))
    hd = float(f32(p.h))
    p.vol_def = f32(float(p.dr) ** 3)
    ih = f32(p.i_h)
    p.cell_num[0] = int(f32(0.5) * ih * f32(p.domain_size[0])) + 1
    p.cell_num[1] = int(f32(0.5) * ih * f32(p.domain_size[1])) + 1
    p.cell_num[2] = int(f32(0.5) * ih * f32(p.domain_size[2])) + 1
    p.kernel_normalization_par = f32(42.0 / (256.0 * math.pi * hd ** 3))
    p.kernel_gradient_normalization_par = f32(-5.0 * 42 / (256.0 * math.pi * hd ** 5))
    layer = p.cell_num[0] * p.cell_num[1] * p.cell_num[2]
    p.ncell_material = layer
    p.ncell_delete_m = 2 * layer
    p.ncell = 2 * layer + 1
    p.np_delete_m = 0
    q = f32(p.dr) * ih
    kn = f32(p.kernel_normalization_par)
    p.kernel_zero = kn * f32(2) * f32(2) * f32(2) * f32(2) * f32(0.5)
    t = f32(2.0) - q
    p.kernel_dr = kn * t * t * t * t * (q + f32(0.5))
    return p


def motion_tables(case: Case):
    arr = (_abi.MotionTable * max(1, len(case.motion)))()
    for k, t in enumerate(case.motion):
        m = arr[k]
        for key in ["is_translation", "table_size_trans", "is_rotation", "pattern", "table_size_rotat"]:
            setattr(m, key, t[key])
        for j in range(10):
            m.trans_vel[j][:] = t["trans_vel"][j]
            m.trans_acc[j][:] = t["trans_acc"][j]
            m.rotat_acc[j][:] = t["rotat_acc"][j]
            m.trans_time[j] = t["trans_time"][j]
            m.rotat_time[j] = t["rotat_time"][j]
        m.pos_rp[:] = t["pos_rp"]
        m.rotat_vel[:] = t["rotat_vel"]
    return arr


# ----------------------------------------------------------- named configs --
def batch_cases(n_cases: int, seed: int, dr: float = 0.005, n_range=(5.0e3, 2.0e5), h_range=(0.1, 0.6),
                **over) -> list:
    """Independent wall cases for a batched context: N_k log-uniform over
    n_range, backfill height H ~ U(h_range), phi_c ~ U(28, 36) deg, e0 ~
    U(0.60, 0.80), translation or toe rotation with equal odds at a wall speed
    ~ U(0.5, 2) mm/s (rotation rate = speed / H), backfill length twice its
    width.  All cases share dr and the solver constants, as a batch needs."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_cases):
        n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        height = float(rng.uniform(*h_range))
        phi = float(rng.uniform(28.0, 36.0))
        e0 = float(rng.uniform(0.60, 0.80))
        speed = float(rng.uniform(0.5e-3, 2.0e-3))
        rotate = bool(rng.integers(0, 2))
        # the soil also fills the bed step (fix_wall_height) and, with no side
        # walls, the wall-thickness margins in y: N = 2W (W + 2bt)(H + fwh) / dr^3
        soil_h = height + over.get("fix_wall_height", 0.04)
        bt = math.ceil(2.0 * over.get("hdr", 1.3)) * dr
        a2 = n * dr ** 3 / soil_h
        width = (-4.0 * bt + math.sqrt(16.0 * bt * bt + 8.0 * a2)) / 4.0
        motion = ("rotate", -speed / height) if rotate else ("translate", -speed)
        kw = dict(box_length=2.0 * width, box_width=width, box_height=height, dr=dr, motion=motion, jitter=0.05,
                  void_range=(e0 - 0.05, e0 + 0.05), seed=seed * 1000 + k,
                  material=hypo_cs_sand(phi_c_deg=phi, e0=e0))
        kw.update(over)
        out.append(retaining_wall(name=f"batch{seed}_{k}", **kw))
    return out


def config(name: str, **over) -> Case:
    """The benchmark configurations (geometry / material / loading)."""
    if name == "c1":  # 0.4 x 0.2 x 0.3 m, dx 0.005, translation away at 1 mm/s
        kw = dict(box_length=0.4, box_width=0.2, box_height=0.3, dr=0.005, motion=("translate", -0.001))
    elif name == "c2":  # 1.0 x 0.5 x 0.6 m at dx 0.0063 (~1.2M), toe rotation (active side)
        kw = dict(box_length=1.0, box_width=0.5, box_height=0.6, dr=0.0063, motion=("rotate", -0.01),
                  jitter=0.05, void_range=(0.65, 0.75), seed=2024)
    elif name == "c2_dx004":  # literal dx 0.004 geometry (~5.0M soil particles)
        kw = dict(box_length=1.0, box_width=0.5, box_height=0.6, dr=0.004, motion=("rotate", -0.01),
                  jitter=0.05, void_range=(0.65, 0.75), seed=2024)
    elif name == "c3":  # 2.0 x 1.0 x 1.0 m at dx 0.004 (~31M), 15 deg slope, 10 kPa strip, translation
        kw = dict(box_length=2.0, box_width=1.0, box_height=1.0, dr=0.004, motion=("translate", -0.001),
                  slope_deg=15.0, surcharge=(1.0e4, 0.5, 1.0), jitter=0.05, seed=3)
    elif name == "tiny":  # parity-test size (~6k particles)
        kw = dict(box_length=0.06, box_width=0.03, box_height=0.05, dr=0.0025, fix_wall_height=0.01,
                  wall_extra=0.01, motion=("translate", -0.001), jitter=0.05, seed=7)
    elif name == "small":  # parity-test size (~30k particles), rotating wall
        kw = dict(box_length=0.1, box_width=0.05, box_height=0.08, dr=0.0025, fix_wall_height=0.01,
                  wall_extra=0.01, motion=("rotate", -0.01), jitter=0.05, void_range=(0.65, 0.75), seed=11)
    else:
        raise KeyError(name)
    kw.update(over)
    return retaining_wall(name=name, **kw)
