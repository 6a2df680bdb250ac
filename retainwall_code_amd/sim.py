"""Python handle over the C ABI (libretainwall.so) for tests and benchmarks.

Mirrors the reference's LoquatSph step pieces (loquat.h:97-120) one method per
C entry point; every call raises on a non-zero status with lq_last_error().
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class LqError(RuntimeError):
    pass


class Context:
    def __init__(self, params: _abi.Params, device: int = 0, lib=None):
        self.lib = lib or _abi.load_library()
        self.params = params
        h = C.c_void_p()
        st = self.lib.lq_create(C.byref(params), device, C.byref(h))
        self.h = h
        if st:
            msg = self.lib.lq_last_error(h).decode() if h.value else ""
            if h.value:
                self.lib.lq_destroy(h)
            self.h = None
            raise LqError(f"lq_create failed ({st}): {msg}")
        if params.key_mode != _abi.KEY_LINEAR:  # lq_create sizes the cell arrays for the key space
            self.params = self.get_params()

    # -------------------------------------------------------------- util --
    def _chk(self, st, what):
        if st:
            raise LqError(f"{what} failed ({st}): {self.lib.lq_last_error(self.h).decode()}")

    def close(self):
        if self.h is not None:
            self.lib.lq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- transfer --
    def upload(self, case):
        self._keep = [np.ascontiguousarray(case.codes, np.int32), np.ascontiguousarray(case.pospres, np.float32),
                      np.ascontiguousarray(case.velrhop, np.float32), np.ascontiguousarray(case.idmassidx, np.float32),
                      np.ascontiguousarray(case.str, np.float32), np.ascontiguousarray(case.stn, np.float32),
                      np.ascontiguousarray(case.normal, np.float32)]
        hs = _abi.HostState(*[a.ctypes.data for a in self._keep])
        self._chk(self.lib.lq_upload(self.h, C.byref(hs)), "lq_upload")

    def set_motion(self, tables, n=None):
        n = len(tables) if n is None else n
        self._chk(self.lib.lq_set_motion_tables(self.h, n, tables), "lq_set_motion_tables")

    def get_motion(self, n):
        t = (_abi.MotionTable * max(n, 1))()
        self._chk(self.lib.lq_get_motion_tables(self.h, n, t), "lq_get_motion_tables")
        return t

    def field_shape(self, name):
        fid, dt, w, rows = _abi.FIELDS[name]
        n = self.params.np if rows == "np" else self.params.ncell
        return (n, w) if w > 1 else (n,), np.dtype(dt), fid

    def get(self, name):
        shape, dt, fid = self.field_shape(name)
        out = np.empty(shape, dt)
        self._chk(self.lib.lq_get_field(self.h, fid, out.ctypes.data, out.nbytes), f"lq_get_field({name})")
        return out

    def set(self, name, value):
        shape, dt, fid = self.field_shape(name)
        a = np.ascontiguousarray(np.asarray(value, dt).reshape(shape))
        self._chk(self.lib.lq_set_field(self.h, fid, a.ctypes.data, a.nbytes), f"lq_set_field({name})")

    def snapshot(self, names=None):
        names = names or [k for k in _abi.FIELDS if k not in ("part_cel", "part_idx")]
        return {k: self.get(k) for k in names}

    def get_params(self):
        p = _abi.Params()
        self._chk(self.lib.lq_get_params(self.h, C.byref(p)), "lq_get_params")
        return p

    def time(self):
        t = C.c_double()
        s = C.c_int64()
        self._chk(self.lib.lq_get_time(self.h, C.byref(t), C.byref(s)), "lq_get_time")
        return t.value, s.value

    def set_time(self, t, step):
        self._chk(self.lib.lq_set_time(self.h, t, step), "lq_set_time")

    def set_surcharge(self, block, pressure):
        """Strip load: slots of soil block `block` carry pressure * dr^2 downward."""
        self._chk(self.lib.lq_set_surcharge(self.h, int(block), float(pressure)), "lq_set_surcharge")

    def checkpoint_write(self, path):
        """Whole device state, raw, for a bit-exact restart (lq_checkpoint_write)."""
        self._chk(self.lib.lq_checkpoint_write(self.h, str(path).encode()), "lq_checkpoint_write")

    def checkpoint_read(self, path):
        self._chk(self.lib.lq_checkpoint_read(self.h, str(path).encode()), "lq_checkpoint_read")

    # -------------------------------------------------------- step pieces --
    def step_calculation(self):
        self._chk(self.lib.lq_step_calculation(self.h), "lq_step_calculation")

    def neighbor_search(self):
        self._chk(self.lib.lq_neighbor_search(self.h), "lq_neighbor_search")

    def sort_arrays_primary(self):
        self._chk(self.lib.lq_sort_arrays_primary(self.h), "lq_sort_arrays_primary")

    def sort_arrays_all(self):
        self._chk(self.lib.lq_sort_arrays_all(self.h), "lq_sort_arrays_all")

    def rigid_body_update(self):
        self._chk(self.lib.lq_rigid_body_update(self.h), "lq_rigid_body_update")

    def adami(self):
        self._chk(self.lib.lq_adami_boundary(self.h), "lq_adami_boundary")

    def find_corner(self):
        self._chk(self.lib.lq_find_corner_particle(self.h), "lq_find_corner_particle")

    def shepard(self):
        self._chk(self.lib.lq_stress_regularization(self.h), "lq_stress_regularization")

    def interaction(self):
        self._chk(self.lib.lq_particle_interaction(self.h), "lq_particle_interaction")

    def predictor(self):
        self._chk(self.lib.lq_predictor(self.h), "lq_predictor")

    def corrector(self):
        self._chk(self.lib.lq_corrector(self.h), "lq_corrector")

    def step(self, n=1, graph=False):
        self._chk(self.lib.lq_step(self.h, n, 1 if graph else 0), "lq_step")

    def sync(self):
        self._chk(self.lib.lq_synchronize(self.h), "lq_synchronize")

    def counts(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.lq_get_counts(self.h, C.byref(a), C.byref(b), C.byref(c)), "lq_get_counts")
        return a.value, b.value, c.value

    def wall_force(self, box):
        b = (C.c_float * 3)(*box)
        F = C.c_double()
        self._chk(self.lib.lq_wall_force(self.h, b, C.byref(F)), "lq_wall_force")
        return F.value

    def wall_pressure_profile(self, geom, nbins):
        """(depth, p1, p2, p3) per depth bin (lq_wall_pressure_profile);
        geom = (x_face, z_lo, z_hi, z_top)."""
        g = (C.c_float * 4)(*geom)
        out = np.zeros((nbins, 4), np.float32)
        self._chk(self.lib.lq_wall_pressure_profile(self.h, g, nbins, out.ctypes.data, out.nbytes),
                  "lq_wall_pressure_profile")
        return out

    def wall_pressure_profile_noalloc(self, geom, nbins):
        """lq_wall_pressure_profile_noalloc: the same rows, no device allocation."""
        g = (C.c_float * 4)(*geom)
        out = np.zeros((nbins, 4), np.float32)
        self._chk(self.lib.lq_wall_pressure_profile_noalloc(self.h, g, nbins, out.ctypes.data, out.nbytes),
                  "lq_wall_pressure_profile_noalloc")
        return out

    def sensor_stress(self, points, reference_rows=False):
        """Shepard-weighted stress at absolute points (lq_sensor_stress):
        f32[n][8] = (xx yy zz xy xz yz, sum W V, slots)."""
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        out = np.zeros((len(pts), 8), np.float32)
        self._chk(self.lib.lq_sensor_stress(self.h, pts.ctypes.data, len(pts), 1 if reference_rows else 0,
                                            out.ctypes.data, out.nbytes), "lq_sensor_stress")
        return out

    def surface_heights(self, points, radius):
        """Largest absolute soil z near the vertical lines through (x, y)
        (lq_surface_heights); -1e10 where none."""
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))
        out = np.zeros(len(pts), np.float32)
        self._chk(self.lib.lq_surface_heights(self.h, pts.ctypes.data, len(pts), C.c_float(radius),
                                              out.ctypes.data, out.nbytes), "lq_surface_heights")
        return out

    def positions_absolute(self):
        out = np.empty((self.params.np, 4), np.float32)
        self._chk(self.lib.lq_positions_absolute(self.h, out.ctypes.data, out.nbytes), "lq_positions_absolute")
        return out

    def enable_timing(self, on=True):
        self._chk(self.lib.lq_enable_timing(self.h, 1 if on else 0), "lq_enable_timing")

    def enable_timing_classes(self, names):
        """Time only the named kernel classes (see _abi.KERNEL_CLASSES); resets the counters."""
        mask = 0
        for n in names:
            mask |= 1 << _abi.KERNEL_CLASSES.index(n)
        self._chk(self.lib.lq_enable_timing_classes(self.h, mask), "lq_enable_timing_classes")

    def stream_copy(self, nbytes, reps=20):
        ms = C.c_double()
        self._chk(self.lib.lq_stream_copy(self.h, nbytes, reps, C.byref(ms)), "lq_stream_copy")
        return ms.value

    def kernel_times(self):
        n = len(_abi.KERNEL_CLASSES)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        self._chk(self.lib.lq_kernel_times(self.h, ms, cnt, n), "lq_kernel_times")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(_abi.KERNEL_CLASSES)}

    def set_tile(self, shape, tz):
        self._chk(self.lib.lq_set_tile(self.h, shape, tz), "lq_set_tile")

    def set_interaction(self, variant, xcd_remap=1):
        self._chk(self.lib.lq_set_interaction(self.h, variant, xcd_remap), "lq_set_interaction")

    def set_columns_build(self, build):
        """COLUMNS register budget (lq_set_columns_build): 0 at most 96 VGPRs, 1 at most 128."""
        self._chk(self.lib.lq_set_columns_build(self.h, int(build)), "lq_set_columns_build")

    def set_overlap(self, on=True):
        """COLUMNS steps: corner search beside the interaction (lq_set_overlap)."""
        self._chk(self.lib.lq_set_overlap(self.h, 1 if on else 0), "lq_set_overlap")

    def set_fusion(self, on=True):
        """Fused COLUMNS steps (lq_set_fusion)."""
        self._chk(self.lib.lq_set_fusion(self.h, 1 if on else 0), "lq_set_fusion")

    def set_pair_list(self, cap=160, skin=0.01):
        self._chk(self.lib.lq_set_pair_list(self.h, cap, skin), "lq_set_pair_list")

    def pair_list_status(self):
        """(overflow, max displacement since the last build, allowed displacement)."""
        ov, md, lim = C.c_int(), C.c_float(), C.c_float()
        self._chk(self.lib.lq_pair_list_status(self.h, C.byref(ov), C.byref(md), C.byref(lim)),
                  "lq_pair_list_status")
        return ov.value, md.value, lim.value

    def autotune(self, reps=3):
        """Time every interaction variant on the current state, keep the
        fastest; returns (variant, tile_shape, tile_tz, xcd_remap, ms_per_call)."""
        v, sh, tz, rm, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._chk(self.lib.lq_autotune(self.h, reps, C.byref(v), C.byref(sh), C.byref(tz), C.byref(rm),
                                       C.byref(ms)), "lq_autotune")
        return v.value, sh.value, tz.value, rm.value, ms.value

    def neighbor_lists(self, max_nb=160):
        npm = self.counts()[1]
        counts = np.zeros(npm, np.int32)
        ids = np.zeros((npm, max_nb), np.int32)
        self._chk(self.lib.lq_neighbor_lists(self.h, max_nb, counts.ctypes.data, ids.ctypes.data, ids.nbytes),
                  "lq_neighbor_lists")
        return counts, ids


class BatchContext(Context):
    """Many independent cases stepped together (lq_create_batch): one launch
    per kernel over all cases; case k's results equal a Context of case k
    alone bit for bit."""

    def __init__(self, params_list, device: int = 0, lib=None):
        self.lib = lib or _abi.load_library()
        self.case_params = list(params_list)
        arr = (_abi.Params * len(self.case_params))(*self.case_params)
        h = C.c_void_p()
        st = self.lib.lq_create_batch(arr, len(self.case_params), device, C.byref(h))
        self.h = h
        if st:
            msg = self.lib.lq_last_error(h).decode() if h.value else ""
            if h.value:
                self.lib.lq_destroy(h)
            self.h = None
            raise LqError(f"lq_create_batch failed ({st}): {msg}")
        self.params = self.get_params()
        self._keep = []

    @property
    def n_cases(self):
        return len(self.case_params)

    def upload_case(self, k, case):
        keep = [np.ascontiguousarray(case.codes, np.int32), np.ascontiguousarray(case.pospres, np.float32),
                np.ascontiguousarray(case.velrhop, np.float32), np.ascontiguousarray(case.idmassidx, np.float32),
                np.ascontiguousarray(case.str, np.float32), np.ascontiguousarray(case.stn, np.float32),
                np.ascontiguousarray(case.normal, np.float32)]
        hs = _abi.HostState(*[a.ctypes.data for a in keep])
        self._chk(self.lib.lq_upload_case(self.h, k, C.byref(hs)), "lq_upload_case")

    def set_case_motion(self, k, tables, n=None):
        n = len(tables) if n is None else n
        self._chk(self.lib.lq_set_case_motion_tables(self.h, k, n, tables), "lq_set_case_motion_tables")

    def get_case(self, k, name):
        fid, dt, w, rows = _abi.FIELDS[name]
        n = self.case_params[k].np
        out = np.empty((n, w) if w > 1 else (n,), np.dtype(dt))
        self._chk(self.lib.lq_get_case_field(self.h, k, fid, out.ctypes.data, out.nbytes), f"lq_get_case_field({name})")
        return out

    def case_snapshot(self, k, names):
        return {n: self.get_case(k, n) for n in names}

    def case_positions_absolute(self, k):
        out = np.empty((self.case_params[k].np, 4), np.float32)
        self._chk(self.lib.lq_case_positions_absolute(self.h, k, out.ctypes.data, out.nbytes),
                  "lq_case_positions_absolute")
        return out

    def case_counts(self, k):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.lq_case_counts(self.h, k, C.byref(a), C.byref(b), C.byref(c)), "lq_case_counts")
        return a.value, b.value, c.value

    def wall_force_cases(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, np.float32).reshape(-1, 3))
        F = np.zeros(len(b), np.float64)
        self._chk(self.lib.lq_wall_force_cases(self.h, b.ctypes.data, F.ctypes.data), "lq_wall_force_cases")
        return F


def open_batch(case_list, key_mode=_abi.KEY_LINEAR, variant=_abi.INTERACTION_NEIGHBOR_LIST, device=0):
    """Batched context over several cases: upload every case, its motion
    tables, then the initial search/sort (loquat.cu:528-531)."""
    from . import cases
    lib = _abi.load_library()
    plist = [cases.make_params(c, key_mode, variant, lib=lib) for c in case_list]
    ctx = BatchContext(plist, device, lib)
    for k, c in enumerate(case_list):
        ctx.upload_case(k, c)
        ctx.set_case_motion(k, cases.motion_tables(c), len(c.motion))
    ctx.neighbor_search()
    ctx.sort_arrays_primary()
    ctx.sync()
    return ctx


def open_case(case, key_mode=_abi.KEY_LINEAR, variant=_abi.INTERACTION_NEIGHBOR_LIST, device=0, initial_sort=True):
    """Create a context, upload a case and run the initial search/sort
    (loquat.cu:528-531: GridConfig, Abs2Rel, NeighborSearch, SortArraysPrimary)."""
    from . import cases
    lib = _abi.load_library()
    p = cases.make_params(case, key_mode, variant, lib=lib)
    ctx = Context(p, device, lib)
    ctx.upload(case)
    ctx.set_motion(cases.motion_tables(case), len(case.motion))
    if case.surcharge is not None:
        ctx.set_surcharge(*case.surcharge)
    if initial_sort:
        ctx.neighbor_search()
        ctx.sort_arrays_primary()
        ctx.sync()
    return ctx
