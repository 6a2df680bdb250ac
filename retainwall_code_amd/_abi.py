"""ctypes mirror of include/retainwall.h (structs, enums, prototypes).

The product library ``libretainwall.so`` is built in-tree next to this file;
``load_library`` raises if it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libretainwall.so")

MODEL_NUMBER_MAX = 10
MATERIAL_SIZE_MAX = 15
RIGID_MOTION_MAX = 10
TABLE_MAX = 10

# lq_status
OK, ERR_ARG, ERR_HIP, ERR_DELETE, ERR_DOMAIN, ERR_STATE = range(6)
# lq_model_mat (types.h:47-50)
RIGID, DP_EPP, MUI_DP_EPP, HYPO_BASIC, HYPO_CS = range(5)
FIX, NONFIX = 0, 1
CUBOID, CIRCLE = 0, 1
KEY_LINEAR, KEY_MORTON = 0, 1
PROFILE_NOALLOC_MAX_BINS = 4096  # LQ_PROFILE_NOALLOC_MAX_BINS
INTERACTION_REFERENCE_ORDER, INTERACTION_TILE, INTERACTION_NEIGHBOR_LIST, INTERACTION_PAIR_LIST = 0, 1, 2, 3
INTERACTION_COLUMNS = 4
MLP_IN, MLP_OUT, MLP_HIDDEN, MLP_NPARAM, MLP_NNORM = 14, 7, 128, 35847, 42

# lq_field
(F_CODES, F_CELLPOS, F_POS, F_POS_ORI, F_VEL, F_IDM, F_POS_INI, F_VEL_INI, F_ACCAD, F_SPIN,
 F_CVAR, F_NORMAL, F_STR, F_STN, F_STR_INI, F_STN_INI, F_ACESTN, F_CMAG, F_PART_CEL,
 F_PART_IDX, F_CELL_BEG, F_CELL_END, F_ORIG_ID, F_COUNT) = range(24)

# name -> (field id, dtype, width, rows: "np" | "ncell")
FIELDS = {
    "codes": (F_CODES, "int32", 4, "np"),
    "cellpos": (F_CELLPOS, "int32", 4, "np"),
    "pos": (F_POS, "float32", 4, "np"),
    "pos_ori": (F_POS_ORI, "float32", 4, "np"),
    "vel": (F_VEL, "float32", 4, "np"),
    "idm": (F_IDM, "float32", 4, "np"),
    "pos_ini": (F_POS_INI, "float32", 4, "np"),
    "vel_ini": (F_VEL_INI, "float32", 4, "np"),
    "accad": (F_ACCAD, "float32", 4, "np"),
    "spin": (F_SPIN, "float32", 4, "np"),
    "cvar": (F_CVAR, "float32", 4, "np"),
    "normal": (F_NORMAL, "float32", 4, "np"),
    "str": (F_STR, "float32", 6, "np"),
    "stn": (F_STN, "float32", 6, "np"),
    "str_ini": (F_STR_INI, "float32", 6, "np"),
    "stn_ini": (F_STN_INI, "float32", 6, "np"),
    "acestn": (F_ACESTN, "float32", 6, "np"),
    "cmag": (F_CMAG, "float32", 9, "np"),
    "part_cel": (F_PART_CEL, "int32", 1, "np"),
    "part_idx": (F_PART_IDX, "int32", 1, "np"),
    "cell_beg": (F_CELL_BEG, "int32", 1, "ncell"),
    "cell_end": (F_CELL_END, "int32", 1, "ncell"),
    "orig_id": (F_ORIG_ID, "int32", 1, "np"),
}

# lq_kernel_class
KERNEL_CLASSES = ["neighbor_search", "sort_arrays", "adami", "corner", "interaction",
                  "contact", "predictor", "corrector", "shepard", "pair_list_build", "column_index"]


class RigidMotion(C.Structure):
    _fields_ = [
        ("is_trans", C.c_int32), ("is_rotat", C.c_int32),
        ("trans_vel", C.c_float * 3), ("trans_acc", C.c_float * 3),
        ("pattern", C.c_int32),
        ("pos_rp", C.c_float * 3), ("rotat_vel", C.c_float * 3), ("rotat_acc", C.c_float * 3),
    ]


class MotionTable(C.Structure):
    _fields_ = [
        ("is_translation", C.c_int32), ("table_size_trans", C.c_int32),
        ("trans_vel", (C.c_float * 3) * TABLE_MAX), ("trans_time", C.c_float * TABLE_MAX),
        ("trans_acc", (C.c_float * 3) * TABLE_MAX),
        ("is_rotation", C.c_int32), ("pattern", C.c_int32), ("table_size_rotat", C.c_int32),
        ("pos_rp", C.c_float * 3), ("rotat_vel", C.c_float * 3),
        ("rotat_time", C.c_float * TABLE_MAX), ("rotat_acc", (C.c_float * 3) * TABLE_MAX),
    ]


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("periodbound", C.c_int32),
        ("planeleft", C.c_float * 4), ("planeright", C.c_float * 4),
        ("n_motion", C.c_int32), ("mv", RigidMotion * RIGID_MOTION_MAX),
        ("np", C.c_int32), ("npb", C.c_int32), ("npm", C.c_int32), ("np_delete_m", C.c_int32),
        ("domain_size", C.c_float * 3), ("domain_min", C.c_float * 3),
        ("cell_num", C.c_int32 * 3),
        ("ncell", C.c_int32), ("ncell_material", C.c_int32), ("ncell_delete_m", C.c_int32),
        ("h", C.c_float), ("i_h", C.c_float), ("dr", C.c_float), ("hdr", C.c_float),
        ("vol_def", C.c_float), ("bound_thickness", C.c_float),
        ("gravity", C.c_float * 3),
        ("rho0_m", C.c_float), ("damping", C.c_float),
        ("n_model", C.c_int32),
        ("model_par", (C.c_float * MATERIAL_SIZE_MAX) * MODEL_NUMBER_MAX),
        ("kernel_normalization_par", C.c_float), ("kernel_gradient_normalization_par", C.c_float),
        ("stress_shepard_filter", C.c_int32), ("dynamic_time_step", C.c_int32),
        ("integration_method", C.c_int32), ("viscosity_type", C.c_int32),
        ("artificial_viscosity_alpha_m", C.c_float), ("cs_m", C.c_float), ("eta", C.c_float),
        ("cfl", C.c_float),
        ("artificial_pressure_coefficient_material", C.c_float), ("kernel_zero", C.c_float),
        ("kernel_dr", C.c_float),
        ("time_max", C.c_float), ("time", C.c_float), ("dt", C.c_float),
        ("output_frequency", C.c_float),
        ("output_number", C.c_int32), ("output_format", C.c_int32),
        ("key_mode", C.c_int32), ("interaction_variant", C.c_int32),
    ]


class HostState(C.Structure):
    _fields_ = [
        ("codes", C.c_void_p), ("pospres", C.c_void_p), ("velrhop", C.c_void_p),
        ("idmassidx", C.c_void_p), ("str", C.c_void_p), ("stn", C.c_void_p), ("normal", C.c_void_p),
    ]


EXPORTS = [
    "lq_abi_sizes", "lq_initialize_parameters", "lq_create", "lq_destroy", "lq_last_error",
    "lq_upload", "lq_set_motion_tables", "lq_get_motion_tables", "lq_get_field", "lq_set_field",
    "lq_get_params", "lq_get_time", "lq_set_time", "lq_checkpoint_write", "lq_checkpoint_read",
    "lq_step_calculation", "lq_neighbor_search",
    "lq_sort_arrays_primary", "lq_sort_arrays_all", "lq_rigid_body_update", "lq_adami_boundary",
    "lq_find_corner_particle", "lq_stress_regularization", "lq_particle_interaction",
    "lq_predictor", "lq_corrector", "lq_step", "lq_wall_force", "lq_wall_pressure_profile", "lq_wall_pressure_profile_noalloc", "lq_set_surcharge", "lq_positions_absolute",
    "lq_get_counts", "lq_synchronize", "lq_stress_saving_regularization", "lq_stream_copy", "lq_enable_timing", "lq_enable_timing_classes", "lq_kernel_times", "lq_neighbor_lists", "lq_set_tile", "lq_set_interaction", "lq_set_fusion", "lq_set_overlap", "lq_set_columns_build", "lq_sensor_stress", "lq_surface_heights", "lq_set_pair_list", "lq_pair_list_status", "lq_autotune", "lq_interaction_cutoff_sq",
    "lq_create_batch", "lq_case_count", "lq_upload_case", "lq_set_case_motion_tables", "lq_get_case_field",
    "lq_case_positions_absolute", "lq_case_counts", "lq_wall_force_cases",
    "lq_mlp_normalisation", "lq_mlp_train", "lq_mlp_loss_grad", "lq_mlp_forward", "lq_mlp_last_error",
    "lq_surrogate_set", "lq_surrogate_enable", "lq_surrogate_stats", "lq_surrogate_log", "lq_surrogate_log_read",
]


def _proto(lib):
    P = C.POINTER
    vp = C.c_void_p
    i32 = C.c_int32
    lib.lq_abi_sizes.argtypes = [P(C.c_size_t), C.c_int]
    lib.lq_initialize_parameters.argtypes = [P(Params)]
    lib.lq_create.argtypes = [P(Params), C.c_int, P(vp)]
    lib.lq_destroy.argtypes = [vp]
    lib.lq_destroy.restype = None
    lib.lq_last_error.argtypes = [vp]
    lib.lq_last_error.restype = C.c_char_p
    lib.lq_upload.argtypes = [vp, P(HostState)]
    lib.lq_set_motion_tables.argtypes = [vp, C.c_int, P(MotionTable)]
    lib.lq_get_motion_tables.argtypes = [vp, C.c_int, P(MotionTable)]
    lib.lq_get_field.argtypes = [vp, C.c_int, vp, C.c_size_t]
    lib.lq_set_field.argtypes = [vp, C.c_int, vp, C.c_size_t]
    lib.lq_get_params.argtypes = [vp, P(Params)]
    lib.lq_get_time.argtypes = [vp, P(C.c_double), P(C.c_int64)]
    lib.lq_set_time.argtypes = [vp, C.c_double, C.c_int64]
    lib.lq_checkpoint_write.argtypes = [vp, C.c_char_p]
    lib.lq_checkpoint_read.argtypes = [vp, C.c_char_p]
    for name in ["lq_step_calculation", "lq_neighbor_search", "lq_sort_arrays_primary",
                 "lq_sort_arrays_all", "lq_rigid_body_update", "lq_adami_boundary",
                 "lq_find_corner_particle", "lq_stress_regularization", "lq_particle_interaction",
                 "lq_predictor", "lq_corrector", "lq_synchronize"]:
        getattr(lib, name).argtypes = [vp]
    lib.lq_step.argtypes = [vp, C.c_int, C.c_int]
    lib.lq_wall_force.argtypes = [vp, P(C.c_float), P(C.c_double)]
    lib.lq_wall_pressure_profile.argtypes = [vp, P(C.c_float), C.c_int, vp, C.c_size_t]
    lib.lq_wall_pressure_profile_noalloc.argtypes = [vp, P(C.c_float), C.c_int, vp, C.c_size_t]
    lib.lq_set_fusion.argtypes = [vp, C.c_int]
    lib.lq_set_overlap.argtypes = [vp, C.c_int]
    lib.lq_set_columns_build.argtypes = [vp, C.c_int]
    lib.lq_sensor_stress.argtypes = [vp, vp, C.c_int, C.c_int, vp, C.c_size_t]
    lib.lq_surface_heights.argtypes = [vp, vp, C.c_int, C.c_float, vp, C.c_size_t]
    lib.lq_set_surcharge.argtypes = [vp, C.c_int, C.c_float]
    lib.lq_positions_absolute.argtypes = [vp, vp, C.c_size_t]
    lib.lq_stress_saving_regularization.argtypes = [vp, vp, C.c_size_t]
    lib.lq_get_counts.argtypes = [vp, P(i32), P(i32), P(i32)]
    lib.lq_enable_timing.argtypes = [vp, C.c_int]
    lib.lq_enable_timing_classes.argtypes = [vp, C.c_uint]
    lib.lq_stream_copy.argtypes = [vp, C.c_size_t, C.c_int, P(C.c_double)]
    lib.lq_kernel_times.argtypes = [vp, P(C.c_double), P(C.c_int64), C.c_int]
    lib.lq_neighbor_lists.argtypes = [vp, C.c_int, vp, vp, C.c_size_t]
    lib.lq_set_tile.argtypes = [vp, C.c_int, C.c_int]
    lib.lq_set_interaction.argtypes = [vp, C.c_int, C.c_int]
    lib.lq_set_pair_list.argtypes = [vp, C.c_int, C.c_float]
    lib.lq_pair_list_status.argtypes = [vp, P(C.c_int), P(C.c_float), P(C.c_float)]
    lib.lq_autotune.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_double)]
    lib.lq_interaction_cutoff_sq.argtypes = [C.c_float, P(C.c_float)]
    lib.lq_create_batch.argtypes = [P(Params), C.c_int, C.c_int, P(vp)]
    lib.lq_case_count.argtypes = [vp]
    lib.lq_upload_case.argtypes = [vp, C.c_int, P(HostState)]
    lib.lq_set_case_motion_tables.argtypes = [vp, C.c_int, C.c_int, P(MotionTable)]
    lib.lq_get_case_field.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    lib.lq_case_positions_absolute.argtypes = [vp, C.c_int, vp, C.c_size_t]
    lib.lq_case_counts.argtypes = [vp, C.c_int, P(i32), P(i32), P(i32)]
    lib.lq_wall_force_cases.argtypes = [vp, vp, vp]
    lib.lq_mlp_normalisation.argtypes = [vp, vp, C.c_size_t, vp]
    lib.lq_mlp_train.argtypes = [vp, vp, C.c_size_t, vp, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_int, vp, vp]
    lib.lq_mlp_loss_grad.argtypes = [vp, vp, vp, vp, C.c_size_t, vp, P(C.c_double)]
    lib.lq_mlp_forward.argtypes = [vp, vp, vp, C.c_size_t, vp, C.c_int]
    lib.lq_mlp_last_error.argtypes = []
    lib.lq_mlp_last_error.restype = C.c_char_p
    lib.lq_surrogate_set.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_float, C.c_float]
    lib.lq_surrogate_enable.argtypes = [vp, C.c_int]
    lib.lq_surrogate_stats.argtypes = [vp, vp, P(C.c_float), C.c_int]
    lib.lq_surrogate_log.argtypes = [vp, C.c_int, C.c_int]
    lib.lq_surrogate_log_read.argtypes = [vp, vp, vp, C.c_int, P(C.c_int)]


_LIB = None


def load_library(path=None):
    """Load libretainwall.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()'). "
            "There is no CPU fallback for the solver.")
    lib = C.CDLL(p)
    _proto(lib)
    sizes = (C.c_size_t * 4)()
    lib.lq_abi_sizes(sizes, 4)
    expect = [C.sizeof(Params), C.sizeof(MotionTable), C.sizeof(RigidMotion), C.sizeof(HostState)]
    if list(sizes) != expect:
        raise RuntimeError(f"ABI layout mismatch: library {list(sizes)} vs binding {expect}")
    if path is None:
        _LIB = lib
    return lib
