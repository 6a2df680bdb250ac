"""ctypes binding of the CPU ORACLE (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The oracle restates the
reference's per-particle functions serially on the host (see
loquat_oracle.hpp for the per-function citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_ARRAYS4 = ["pos", "pos_ori", "vel", "idm", "pos_ini", "vel_ini", "accad", "spin", "cvar", "normal"]
_ARRAYS6 = ["str", "stn", "str_ini", "stn_ini", "acestn"]
_INT4 = ["codes", "cellpos"]
_INT1 = ["part_cel", "part_idx", "orig_id"]


class OrState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in _ARRAYS4 + _ARRAYS6 + ["cmag"]] + \
               [(n, C.c_void_p) for n in _INT4 + _INT1 + ["cell_beg", "cell_end", "deferred"]]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        from retainwall_code_amd import _abi
        P = C.POINTER
        L.or_initialize_parameters.argtypes = [P(_abi.Params)]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = None
        L.or_set_surcharge.argtypes = [C.c_int, C.c_double]
        L.or_set_surcharge.restype = None
        for fn in ("or_kernel_f64", "or_grad_f64"):
            getattr(L, fn).argtypes = [C.c_double, C.c_double]
            getattr(L, fn).restype = C.c_double
        for sfx, T in (("f32", C.c_float), ("f64", C.c_double)):
            for name in ["abs2rel", "sort_arrays_primary", "sort_arrays_all", "adami", "find_corner",
                         "shepard", "interaction", "predictor", "corrector"]:
                getattr(L, f"or_{name}_{sfx}").argtypes = [P(_abi.Params), P(OrState)]
            getattr(L, f"or_neighbor_search_{sfx}").argtypes = [P(_abi.Params), P(OrState), P(C.c_int)]
            getattr(L, f"or_rigid_body_update_{sfx}").argtypes = [P(_abi.Params), P(_abi.MotionTable)]
            getattr(L, f"or_hypo_cs_{sfx}").argtypes = [T, T, P(T), P(T), P(T), P(T), P(T)]
            getattr(L, f"or_constitutive_{sfx}").argtypes = [C.c_int, T, T, P(T), P(T), P(T), P(T), P(T),
                                                             P(C.c_int)]
            getattr(L, f"or_wall_force_{sfx}").argtypes = [P(_abi.Params), P(OrState), P(C.c_float)]
            getattr(L, f"or_wall_force_{sfx}").restype = C.c_double
            getattr(L, f"or_wall_profile_{sfx}").argtypes = [P(_abi.Params), P(OrState), P(C.c_float), C.c_int,
                                                             C.c_void_p]
            getattr(L, f"or_sensor_stress_{sfx}").argtypes = [P(_abi.Params), P(OrState), C.c_void_p, C.c_int,
                                                              C.c_int, C.c_void_p]
            getattr(L, f"or_surface_heights_{sfx}").argtypes = [P(_abi.Params), P(OrState), C.c_void_p, C.c_int,
                                                                C.c_float, C.c_void_p]
            getattr(L, f"or_neighbor_lists_{sfx}").argtypes = [P(_abi.Params), P(OrState), C.c_int, C.c_void_p,
                                                               C.c_void_p]
            getattr(L, f"or_run_{sfx}").argtypes = [P(_abi.Params), P(_abi.MotionTable), P(OrState), C.c_int,
                                                    P(C.c_double), P(C.c_int64), P(C.c_float), C.c_int,
                                                    P(C.c_double), P(C.c_int), P(C.c_double)]
        _LIB = L
    return _LIB


class Oracle:
    """Host state (numpy) + oracle kernels for one precision."""

    def __init__(self, params, dtype=np.float64, motion=None):
        from retainwall_code_amd import _abi
        self.p = _abi.Params()
        C.memmove(C.byref(self.p), C.byref(params), C.sizeof(_abi.Params))
        self.dtype = np.dtype(dtype)
        self.sfx = "f64" if self.dtype == np.float64 else "f32"
        n = params.np
        self.a = {}
        for k in _ARRAYS4:
            self.a[k] = np.zeros((n, 4), self.dtype)
        for k in _ARRAYS6:
            self.a[k] = np.zeros((n, 6), self.dtype)
        self.a["cmag"] = np.zeros((n, 9), self.dtype)
        for k in _INT4:
            self.a[k] = np.zeros((n, 4), np.int32)
        for k in _INT1 + ["deferred"]:
            self.a[k] = np.zeros(n, np.int32)
        self.a["cell_beg"] = np.full(params.ncell, -1, np.int32)
        self.a["cell_end"] = np.full(params.ncell, -1, np.int32)
        self.tables = motion
        self.time = float(params.time)
        self.step = 0

    def __getitem__(self, k):
        return self.a[k]

    def _st(self):
        s = OrState()
        for name, _ in OrState._fields_:
            arr = self.a[name]
            assert arr.flags["C_CONTIGUOUS"]
            setattr(s, name, arr.ctypes.data)
        return s

    def _call(self, name, *extra):
        s = self._st()
        fn = getattr(lib(), f"or_{name}_{self.sfx}")
        return fn(C.byref(self.p), C.byref(s), *extra)

    # ---- loading ----
    def load_host(self, case):
        """Initial state from a generated case (absolute positions)."""
        n = case.np
        self.a["codes"][:] = case.codes
        self.a["pos"][:] = case.pospres
        self.a["pos_ori"][:] = case.pospres
        self.a["vel"][:] = case.velrhop
        self.a["idm"][:] = case.idmassidx
        self.a["idm"][:, 2] = np.arange(n)
        self.a["str"][:] = case.str
        self.a["stn"][:] = case.stn
        self.a["normal"][:case.npb] = case.normal
        self.a["orig_id"][:] = np.arange(n, dtype=np.int32)
        return self

    def load_fields(self, fields: dict):
        """Copy a state snapshot (e.g. downloaded from the device)."""
        for k, v in fields.items():
            if k in self.a:
                self.a[k][...] = np.asarray(v).astype(self.a[k].dtype).reshape(self.a[k].shape)
        return self

    # ---- kernels ----
    def abs2rel(self):
        return self._call("abs2rel")

    def neighbor_search(self):
        ident = C.c_int(0)
        st = self._call("neighbor_search", C.byref(ident))
        if st:
            raise RuntimeError(f"oracle neighbor_search status {st}")
        return bool(ident.value)

    def sort_arrays_primary(self):
        self._call("sort_arrays_primary")

    def sort_arrays_all(self):
        self._call("sort_arrays_all")

    def rigid_body_update(self):
        getattr(lib(), f"or_rigid_body_update_{self.sfx}")(C.byref(self.p), self.tables)

    def adami(self):
        self._call("adami")

    def find_corner(self):
        self._call("find_corner")

    def shepard(self):
        self._call("shepard")

    def interaction(self):
        self._call("interaction")

    def predictor(self):
        self._call("predictor")

    def corrector(self):
        return self._call("corrector")

    def neighbor_lists(self, max_nb=160):
        npm = self.p.npm
        counts = np.zeros(npm, np.int32)
        ids = np.zeros((npm, max_nb), np.int32)
        self._call("neighbor_lists", max_nb, counts.ctypes.data, ids.ctypes.data)
        return counts, ids

    def wall_force(self, box):
        s = self._st()
        b = (C.c_float * 3)(*box)
        return getattr(lib(), f"or_wall_force_{self.sfx}")(C.byref(self.p), C.byref(s), b)

    def wall_profile(self, geom, nbins):
        """cuGetValuesAlongLine restated (utilities.cu:266-320)."""
        s = self._st()
        g = (C.c_float * 4)(*geom)
        out = np.zeros((nbins, 4), np.float64)
        getattr(lib(), f"or_wall_profile_{self.sfx}")(C.byref(self.p), C.byref(s), g, int(nbins), out.ctypes.data)
        return out

    def sensor_stress(self, points, reference_rows=False):
        """GetSoilStress restated (utilities.cu:470-528) at absolute points:
        [n][8] = (xx yy zz xy xz yz, sum W V, slots)."""
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        out = np.zeros((len(pts), 8), np.float64)
        s = self._st()
        getattr(lib(), f"or_sensor_stress_{self.sfx}")(C.byref(self.p), C.byref(s), pts.ctypes.data, len(pts),
                                                       1 if reference_rows else 0, out.ctypes.data)
        return out

    def surface_heights(self, points, radius):
        """GetSurface / GetSurfaceAtPoint restated (utilities.cu:26-91,
        371-427) on absolute positions at (x, y) points."""
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))
        out = np.zeros(len(pts), np.float64)
        s = self._st()
        getattr(lib(), f"or_surface_heights_{self.sfx}")(C.byref(self.p), C.byref(s), pts.ctypes.data, len(pts),
                                                         C.c_float(radius), out.ctypes.data)
        return out

    def run(self, nsteps, box=(0, 0, 0), force_every=0):
        s = self._st()
        t = C.c_double(self.time)
        st = C.c_int64(self.step)
        nf = nsteps // force_every if force_every else 0
        forces = (C.c_double * max(nf, 1))()
        ident = C.c_int(0)
        secs = C.c_double(0)
        b = (C.c_float * 3)(*box)
        r = getattr(lib(), f"or_run_{self.sfx}")(C.byref(self.p), self.tables, C.byref(s), nsteps, C.byref(t),
                                                C.byref(st), b, force_every, forces, C.byref(ident), C.byref(secs))
        if r:
            raise RuntimeError(f"oracle run status {r}")
        self.time, self.step = t.value, st.value
        return {"forces": np.array(forces[:nf]), "ns2_identity": bool(ident.value), "seconds": secs.value}


def hypo_cs(dt, rho, w, D, sig, mp, dtype=np.float64):
    """Single HypoCS update (constitutive_model.h:209-299)."""
    T = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    sfx = "f64" if T is C.c_double else "f32"
    wv = (T * 3)(*w)
    Dv = (T * 6)(*D)
    sv = (T * 6)(*sig)
    mv = (T * 15)(*mp)
    mc = T(0)
    getattr(lib(), f"or_hypo_cs_{sfx}")(T(dt), T(rho), wv, Dv, sv, mv, C.byref(mc))
    return np.array(sv[:], dtype=dtype), mc.value


def constitutive(model_type, dt, rho, w, D, sig, mp, dtype=np.float64):
    """One stress update by the model dispatch of cuPredictorMaterial
    (predictor_corrector.cu:72-75): DP, mu(I)-DP, basic or critical-state
    hypoplasticity.  Returns (sigma, M_crt, yielded)."""
    T = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    sfx = "f64" if T is C.c_double else "f32"
    wv = (T * 3)(*w)
    Dv = (T * 6)(*D)
    sv = (T * 6)(*sig)
    mv = (T * 15)(*mp)
    mc = T(0)
    y = C.c_int(0)
    getattr(lib(), f"or_constitutive_{sfx}")(int(model_type), T(dt), T(rho), wv, Dv, sv, mv, C.byref(mc), C.byref(y))
    return np.array(sv[:], dtype=dtype), mc.value, bool(y.value)


def set_threads(n):
    """OpenMP threads for the oracle's per-particle loops (results are the
    same for any count)."""
    lib().or_set_threads(int(n))


def set_surcharge(block, pressure):
    """Strip surcharge for the following oracle calls: slots of soil block
    `block` carry pressure * dr^2 downward (block -1: none)."""
    lib().or_set_surcharge(int(block), float(pressure))


def kernel(knorm, q):
    return lib().or_kernel_f64(knorm, q)


def grad(kgrad, q):
    return lib().or_grad_f64(kgrad, q)


def initialize_parameters(p):
    return lib().or_initialize_parameters(C.byref(p))
