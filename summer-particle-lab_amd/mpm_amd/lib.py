"""ctypes binding of the engine's C ABI (include/mpm_engine.h).

Loads ``lib/libmpm_amd.so`` from this package's parent directory. There is no
fallback: if the library is missing the import of the engine fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmpm_amd.so")

MPM_OK = 0
MPM_ERR_INVALID_ARGUMENT = 1
MPM_ERR_HIP = 2
MPM_ERR_CAPACITY = 3
MPM_ERR_UNSUPPORTED = 4
MPM_ERR_NOT_READY = 5
MPM_ERR_CFL = 7
MPM_ERR_INVALID_STATE = 8

MPM_WALL_SEPARATE, MPM_WALL_STICKY, MPM_WALL_SLIP = 0, 1, 2
MPM_F64, MPM_F32 = 0, 1


class MpmConfig(C.Structure):
    _fields_ = [
        ("lower", C.c_double * 3), ("upper", C.c_double * 3), ("cell_size", C.c_double),
        ("capacity", C.c_uint32), ("wall_mode", C.c_int32 * 6), ("gravity", C.c_double * 3),
        ("gravity_on_particles", C.c_int32), ("mass_eps", C.c_double), ("sort_every", C.c_int32),
        ("strict_domain", C.c_int32), ("guard_off", C.c_int32), ("cfl_number", C.c_double),
        ("cfl_cap", C.c_double), ("grad_number", C.c_double),
    ]


_MAT_FLOATS = ["E", "mu", "lambda_",
               "th_Tref", "th_Thot", "th_hotRatio", "th_A", "th_H", "th_rate", "th_Q", "th_n", "th_stress",
               "th_kelvin_offset",
               "pl_A", "pl_B", "pl_n", "pl_C", "pl_rate0", "pl_Tref", "pl_Tm", "pl_m",
               "yield_strain", "friction",
               "pr_cohesion", "pr_slope", "pr_dilatancy", "pr_tension", "pr_alpha", "pr_plateau", "pr_gamma",
               "pr_densify", "pr_beta", "pr_epsD"]


class MpmMaterial(C.Structure):
    _fields_ = ([("model", C.c_int32)] + [(k, C.c_float) for k in _MAT_FLOATS] +
                [("fracture", C.c_int32), ("fr_onset", C.c_float), ("fr_failure", C.c_float)])


class MpmParticles(C.Structure):
    _fields_ = [
        ("count", C.c_uint32), ("dtype", C.c_int32),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
        ("F", C.c_void_p), ("C", C.c_void_p),
        ("mass", C.c_void_p), ("volume", C.c_void_p), ("temperature", C.c_void_p),
        ("plastic", C.c_void_p), ("cached_tau", C.c_void_p),
        ("material", C.c_void_p), ("pinned", C.c_void_p), ("cache_valid", C.c_void_p),
        ("jp", C.c_void_p), ("dilation", C.c_void_p), ("damage", C.c_void_p),
    ]


class MpmSnapshotStats(C.Structure):
    _fields_ = [("count", C.c_uint32), ("render_count", C.c_uint32), ("sample_stride", C.c_uint32),
                ("solids", C.c_int64), ("liquids", C.c_int64), ("models", C.c_int64 * 12),
                ("mass", C.c_double), ("min_temperature", C.c_double), ("max_temperature", C.c_double),
                ("min_J", C.c_double), ("max_stress", C.c_double), ("max_plastic", C.c_double),
                ("nonfinite", C.c_int64)]


class Mpm2Particles(C.Structure):
    _fields_ = [("count", C.c_uint32), ("dtype", C.c_int32), ("x", C.c_void_p), ("y", C.c_void_p),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("F", C.c_void_p), ("C", C.c_void_p),
                ("mass", C.c_void_p), ("volume", C.c_void_p), ("pinned", C.c_void_p)]


class Mpm2Settings(C.Structure):
    _fields_ = [("stiffness", C.c_float), ("walls", C.c_int32), ("bound", C.c_float), ("floor", C.c_float),
                ("top", C.c_float)]


class Mpm2Stats(C.Structure):
    _fields_ = [("count", C.c_uint32), ("min_J", C.c_double), ("max_J", C.c_double), ("strain", C.c_double),
                ("nodes", C.c_uint32), ("substeps", C.c_uint32), ("grid_mass_q", C.c_int64),
                ("particle_mass_q", C.c_int64)]


class MpmConservation(C.Structure):
    _fields_ = [
        ("particle_mass", C.c_double), ("grid_mass", C.c_double),
        ("particle_mass_q", C.c_int64), ("grid_mass_q", C.c_int64), ("mass_scale", C.c_double),
        ("momentum", C.c_double * 3), ("grid_momentum", C.c_double * 3), ("angular", C.c_double * 3),
        ("kinetic", C.c_double), ("affine", C.c_double), ("min_J", C.c_double), ("max_speed", C.c_double),
        ("nonfinite", C.c_int64),
    ]


class MpmMlpTrainConfig(C.Structure):
    _fields_ = [("samples", C.c_uint32), ("batch", C.c_uint32), ("epochs", C.c_uint32), ("seed", C.c_uint32),
                ("lr", C.c_float), ("lr_final", C.c_float), ("strain_sigma", C.c_float), ("T_lo", C.c_float),
                ("T_hi", C.c_float), ("dt", C.c_float), ("eps_max", C.c_float), ("near_yield_fraction", C.c_float)]


class MpmMlpNorm(C.Structure):
    _fields_ = ([(k, C.c_float) for k in ("in_scale", "T_center", "T_scale", "out_scale", "eps_max", "T_lo", "T_hi",
                                          "min_yield_ratio")]
                + [("material", C.c_int32), ("model", C.c_int32)]
                + [(k, C.c_float) for k in ("mu_over_E", "lambda_over_E", "A", "Tref", "Thot", "hotRatio")])


MLP_NPARAM = 9222


class MpmBatchOptions(C.Structure):
    _fields_ = [("sort_every", C.c_int32), ("use_graph", C.c_int32), ("particle_blocks", C.c_int32)]


class MpmBatchTuning(C.Structure):
    _fields_ = [("buckets", C.c_int32), ("bucket_edge", C.c_int32 * 4), ("bucket_scenes", C.c_int32 * 4),
                ("particles_per_group", C.c_int32 * 4), ("sort_every", C.c_int32), ("particle_blocks", C.c_int32),
                ("ms_per_step", C.c_double), ("bucket_ms", C.c_double * 4)]


class MpmFrameConfig(C.Structure):
    _fields_ = [("time_step_scale", C.c_double), ("max_substeps", C.c_int32), ("velocity_limit", C.c_double),
                ("bounds_lo", C.c_double * 3), ("bounds_hi", C.c_double * 3), ("anti_clump", C.c_int32),
                ("volume_unit", C.c_double), ("gravity_mode", C.c_int32), ("gravity_strength", C.c_double),
                ("gravity_direction", C.c_double * 3), ("gravity_center", C.c_double * 3),
                ("gravity_softening", C.c_double)]


class MpmFrameReport(C.Structure):
    _fields_ = [("substeps", C.c_int32), ("advanced", C.c_double)]


class MpmKernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double)]


# Every symbol declared in include/mpm_engine.h, with (restype, argtypes).
_P = C.c_void_p
SIGNATURES = {
    "mpm_create": (C.c_int, [C.POINTER(MpmConfig), C.POINTER(_P)]),
    "mpm_destroy": (C.c_int, [_P]),
    "mpm_last_error": (C.c_char_p, [_P]),
    "mpm_set_materials": (C.c_int, [_P, C.POINTER(MpmMaterial), C.c_int32]),
    "mpm_upload": (C.c_int, [_P, C.POINTER(MpmParticles)]),
    "mpm_download": (C.c_int, [_P, C.POINTER(MpmParticles)]),
    "mpm_step": (C.c_int, [_P, C.c_double, C.c_int32]),
    "mpm_stable_dt": (C.c_int, [_P, C.c_double, C.c_double, C.POINTER(C.c_double)]),
    "mpm_conservation_report": (C.c_int, [_P, C.POINTER(MpmConservation)]),
    "mpm_refresh_diagnostics": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mpm_stress_tensors": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mpm_preview_layout": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mpm_step_to_grid": (C.c_int, [_P, C.c_double]),
    "mpm_grid_download": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint8)]),
    "mpm_grid_upload_velocity": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mpm_step_from_grid": (C.c_int, [_P, C.c_double, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_double),
                                     C.c_int32]),
    "mpm2_create": (C.c_int, [C.c_uint32, C.POINTER(_P)]),
    "mpm2_destroy": (C.c_int, [_P]),
    "mpm2_last_error": (C.c_char_p, [_P]),
    "mpm2_settings_defaults": (C.c_int, [C.c_char_p, C.POINTER(Mpm2Settings)]),
    "mpm2_upload": (C.c_int, [_P, C.POINTER(Mpm2Particles)]),
    "mpm2_download": (C.c_int, [_P, C.POINTER(Mpm2Particles)]),
    "mpm2_step": (C.c_int, [_P, C.POINTER(Mpm2Settings), C.c_double, C.c_int32]),
    "mpm2_stable_dt": (C.c_int, [_P, C.POINTER(Mpm2Settings), C.POINTER(C.c_double)]),
    "mpm2_advance": (C.c_int, [_P, C.POINTER(Mpm2Settings), C.c_double, C.c_int32, C.c_float, C.c_int32,
                               C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "mpm2_statistics": (C.c_int, [_P, C.POINTER(Mpm2Stats)]),
    "mpm2_synchronize": (C.c_int, [_P]),
    "mpm_snapshot": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_float),
                               C.POINTER(MpmSnapshotStats)]),
    "mpm_frame_packet": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint64,
                                   C.POINTER(C.c_uint8), C.c_uint64, C.POINTER(C.c_uint64)]),
    "mpm_synchronize": (C.c_int, [_P]),
    "mpm_count": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "mpm_frame_defaults": (C.c_int, [_P, C.POINTER(MpmFrameConfig)]),
    "mpm_advance_frame": (C.c_int, [_P, C.POINTER(MpmFrameConfig), C.c_double, C.POINTER(MpmFrameReport)]),
    "mpm_profile": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "mpm_kernel_times": (C.c_int, [_P, C.POINTER(MpmKernelTime), C.c_int32, C.POINTER(C.c_int32)]),
    "mpm_debug_bin": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mpm_sort_pairs": (C.c_int, [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "mpm_sort_pairs_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mpm_debug_p2g": (C.c_int, [_P, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_int32)]),
    "mpm_return_map_batch": (C.c_int, [C.POINTER(MpmMaterial), C.c_uint32, C.c_float] + [C.POINTER(C.c_float)] * 6),
    "mpm_stress_batch": (C.c_int, [C.POINTER(MpmMaterial), C.c_uint32, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float)]),
    "mpm_return_map_state_batch": (C.c_int, [C.POINTER(MpmMaterial), C.c_uint32, C.c_float] +
                                   [C.POINTER(C.c_float)] * 10),
    "mpm_material_defaults": (C.c_int, [C.c_int32, C.c_float, C.POINTER(MpmMaterial)]),
    "mpm_material_check": (C.c_int, [C.POINTER(MpmMaterial)]),
    "mpm_mlp_default_norm": (C.c_int, [C.POINTER(MpmMaterial), C.POINTER(MpmMlpTrainConfig),
                                       C.POINTER(MpmMlpNorm)]),
    "mpm_mlp_check_material": (C.c_int, [C.POINTER(MpmMaterial), C.POINTER(MpmMlpNorm)]),
    "mpm_mlp_samples": (C.c_int, [C.POINTER(MpmMaterial), C.POINTER(MpmMlpTrainConfig), C.POINTER(MpmMlpNorm),
                                  C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mpm_mlp_train": (C.c_int, [C.POINTER(MpmMaterial), C.POINTER(MpmMlpTrainConfig), C.POINTER(MpmMlpNorm),
                                C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mpm_mlp_loss_grad": (C.c_int, [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mpm_mlp_infer": (C.c_int, [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mpm_surrogate_return_batch": (C.c_int, [C.POINTER(MpmMaterial), C.POINTER(C.c_float), C.POINTER(MpmMlpNorm),
                                             C.c_uint32, C.c_float] + [C.POINTER(C.c_float)] * 6
                                   + [C.POINTER(C.c_int32)]),
    "mpm_batch_create": (C.c_int, [C.POINTER(MpmConfig), C.POINTER(C.c_uint32), C.c_int32,
                                   C.POINTER(MpmBatchOptions), C.POINTER(_P)]),
    "mpm_batch_destroy": (C.c_int, [_P]),
    "mpm_batch_last_error": (C.c_char_p, [_P]),
    "mpm_batch_scene_count": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]),
    "mpm_batch_set_materials": (C.c_int, [_P, C.POINTER(MpmMaterial), C.c_int32]),
    "mpm_batch_upload": (C.c_int, [_P, C.c_int32, C.POINTER(MpmParticles)]),
    "mpm_batch_download": (C.c_int, [_P, C.c_int32, C.POINTER(MpmParticles)]),
    "mpm_batch_set_dt": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mpm_batch_step": (C.c_int, [_P, C.c_int32]),
    "mpm_batch_synchronize": (C.c_int, [_P]),
    "mpm_batch_conservation": (C.c_int, [_P, C.c_int32, C.POINTER(MpmConservation)]),
    "mpm_batch_debug_bin": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mpm_batch_get_tuning": (C.c_int, [_P, C.POINTER(MpmBatchTuning)]),
    "mpm_batch_set_tuning": (C.c_int, [_P, C.POINTER(MpmBatchTuning)]),
    "mpm_batch_tune": (C.c_int, [_P, C.c_int32, C.POINTER(MpmBatchTuning)]),
    "mpm_set_surrogate": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_float), C.POINTER(MpmMlpNorm)]),
    "mpm_surrogate_counts": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.c_int32]),
}

_lib = None


class MpmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"MPM error {code}: {message}")
        self.code = code


def load():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"MI355X engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, handle=None, batch=None):
    if rc != MPM_OK:
        if handle:
            msg = load().mpm_last_error(handle).decode()
        elif batch:
            msg = load().mpm_batch_last_error(batch).decode()
        else:
            msg = ""
        raise MpmError(rc, msg)
