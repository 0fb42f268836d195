"""ctypes binding of the engine's C ABI (include/lmc_capi.h).

Loading fails loudly when liblmc.so is missing: there is no Python or CPU implementation of any
compute entry point behind these wrappers.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblmc.so")

ENERGY, E0, QUARTIC, PAIRSITE = 0, 1, 2, 3
LMC_OK = 0

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u64 = C.c_uint64
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_i8p = C.POINTER(C.c_int8)


class KmcReplica(C.Structure):
    _fields_ = [("temperature", C.c_double), ("seed", C.c_uint64), ("replica_id", C.c_uint32),
                ("restart_steps", C.c_uint64), ("restart_energy", C.c_double),
                ("restart_time", C.c_double), ("maximum_steps", C.c_uint64),
                ("rate_corrector", C.c_int32), ("early_stop", C.c_int32),
                ("solute_disp", C.c_int32)]


class CmcReplica(C.Structure):
    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("seed", C.c_uint64),
                ("replica_id", C.c_uint32), ("restart_steps", C.c_uint64),
                ("restart_energy", C.c_double), ("maximum_steps", C.c_uint64)]


# name -> (restype, argtypes); the list of exported symbols checked by the CPU tests.
SIGNATURES = {
    "lmc_last_error": (C.c_char_p, []),
    "lmc_version": (C.c_int, []),
    "lmc_device_alloc": (C.c_int, [_u64, C.POINTER(_vp)]),
    "lmc_device_free": (C.c_int, [_vp]),
    "lmc_copy_to_device": (C.c_int, [_vp, _vp, _u64]),
    "lmc_copy_to_host": (C.c_int, [_vp, _vp, _u64]),
    "lmc_synchronize": (C.c_int, [_vp]),
    "lmc_model_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), _i32, _i32,
                                 C.POINTER(_vp)]),
    "lmc_model_free": (None, [_vp]),
    "lmc_model_num_species": (_i32, [_vp]),
    "lmc_model_species_code": (_i32, [_vp, C.c_char_p]),
    "lmc_model_species_name": (C.c_char_p, [_vp, _i32]),
    "lmc_model_set_formulation": (C.c_int, [_vp, _i32]),
    "lmc_site_change_de": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_total_encode": (C.c_int, [_vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_cluster_energy": (C.c_int, [_vp, _vp, _i32, _i64p, _i32, _f64p, _f64p]),
    "lmc_chemical_potential": (C.c_int, [_vp, C.c_char_p, _f64p]),
    "lmc_config_generate": (C.c_int, [_i32, C.c_char_p, C.POINTER(_vp)]),
    "lmc_config_read": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "lmc_config_read_map": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "lmc_config_free": (None, [_vp]),
    "lmc_config_info": (C.c_int, [_vp, _i32p, _i64p, _f64p]),
    "lmc_config_codes": (C.c_int, [_vp, _i32, C.POINTER(C.c_char_p), _u8p]),
    "lmc_config_atoms": (C.c_int, [_vp, _i64p, _i32p]),
    "lmc_config_update": (C.c_int, [_vp, _i32, C.POINTER(C.c_char_p), _u8p, _i64p, _i32p]),
    "lmc_config_set_basis": (C.c_int, [_vp, _f64p]),
    "lmc_config_write": (C.c_int, [_vp, C.c_char_p]),
    "lmc_config_write_xyz": (C.c_int, [_vp, C.c_char_p]),
    "lmc_config_write_map": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p]),
    "lmc_model_layout": (C.c_int, [_vp, _i32p, _i32p, _i32p]),
    "lmc_model_eci_key": (C.c_int, [_vp, _i32, _i32p, _i32p]),
    "lmc_model_orbits": (C.c_int, [_vp, _i32, _i32, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]),
    "lmc_philox4x32": (None, [_u64, _u64, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "lmc_fcc_num_sites": (_i64, [_i32]),
    "lmc_fcc_neighbors": (_i32, [_i32, _i64, _i32, _i64p]),
    "lmc_jump_state_clusters": (_i32, [_i32, _i64, _i64, _i64p, _i32]),
    "lmc_jump_environment": (_i32, [_i32, _i64, _i64, _i64p]),
    "lmc_jump_symmetries": (C.c_int, [_i32p]),
    "lmc_jump_orbit_clusters": (_i32, [_i32, _i64, _i64, _i32, _i64p, _i32]),
    "lmc_eval_jumps": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "lmc_vacancy_scan": (C.c_int, [_vp, _vp, _i32, _i64, _i32, _vp, _i32, _vp, _vp, _vp, _vp]),
    "lmc_eval_pair": (C.c_int, [_vp, _vp, _i32, _i64, _i64, _f64p, _f64p]),
    "lmc_swap_de": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_site_de": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_total_energy": (C.c_int, [_vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_total_energy_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lmc_check_errors": (C.c_int, [_vp]),
    "lmc_kmc_create": (C.c_int, [_vp, _vp, _i32, _i32p, _u8p, C.POINTER(KmcReplica),
                                 C.POINTER(_vp)]),
    "lmc_kmc_free": (None, [_vp]),
    "lmc_kmc_set_schedule": (C.c_int, [_vp, _i32, _f64p, _f64p]),
    "lmc_kmc_set_trace": (C.c_int, [_vp, _u64]),
    "lmc_kmc_set_formulation": (C.c_int, [_vp, _i32]),
    "lmc_kmc_autotune": (C.c_int, [_vp, _i32, _i32p, _f64p, _f64p]),
    "lmc_kmc_autotune_shapes": (C.c_int, [_vp, _i32, _vp, _vp]),
    "lmc_kmc_run": (C.c_int, [_vp, _u64, _vp]),
    "lmc_kmc_state": (C.c_int, [_vp, _u64p, _f64p, _f64p, _f64p, _f64p, _i64p, _i64p, _f64p,
                                _i32p]),
    "lmc_kmc_escape_info": (C.c_int, [_vp, _f64p, _u64p, _i64p]),
    "lmc_kmc_occupancy": (C.c_int, [_vp, _u8p]),
    "lmc_kmc_trace": (C.c_int, [_vp, _i64p, _f64p, _f64p, _f64p]),
    "lmc_kmc_device_views": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                       C.POINTER(_vp)]),
    "lmc_cmc_create": (C.c_int, [_vp, _vp, _i32, _i32p, _u8p, C.POINTER(CmcReplica),
                                 C.POINTER(_vp)]),
    "lmc_cmc_free": (None, [_vp]),
    "lmc_cmc_set_trace": (C.c_int, [_vp, _u64]),
    "lmc_cmc_set_parallel_trials": (C.c_int, [_vp, _i32]),
    "lmc_cmc_round_left": (C.c_int, [_vp, _i32p]),
    "lmc_cmc_run": (C.c_int, [_vp, _u64, _vp]),
    "lmc_cmc_set_annealing": (C.c_int, [_vp, _i32]),
    "lmc_kmc_set_atoms": (C.c_int, [_vp, _i64p, _i32p]),
    "lmc_kmc_atoms": (C.c_int, [_vp, _i64p, _i32p]),
    "lmc_kmc_set_log": (C.c_int, [_vp, _u64, _u64, _f64p]),
    "lmc_kmc_write_log": (C.c_int, [_vp, C.POINTER(C.c_char_p)]),
    "lmc_cmc_set_log": (C.c_int, [_vp, _u64, _u64, _u64]),
    "lmc_cmc_write_log": (C.c_int, [_vp, C.POINTER(C.c_char_p)]),
    "lmc_cmc_set_atoms": (C.c_int, [_vp, _i64p, _i32p]),
    "lmc_cmc_atoms": (C.c_int, [_vp, _i64p, _i32p]),
    "lmc_cmc_set_lds_occupancy": (C.c_int, [_vp, _i32]),
    "lmc_cmc_annealing_state": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lmc_cmc_state": (C.c_int, [_vp, _u64p, _f64p, _f64p, _u64p]),
    "lmc_cmc_occupancy": (C.c_int, [_vp, _u8p]),
    "lmc_cmc_trace": (C.c_int, [_vp, _i64p, _i64p, _f64p, _i8p]),
    "lmc_surrogate_create": (C.c_int, [_vp, _i32, _u64, C.POINTER(_vp)]),
    "lmc_surrogate_free": (None, [_vp]),
    "lmc_surrogate_layout": (C.c_int, [_vp, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]),
    "lmc_surrogate_get_params": (C.c_int, [_vp, _f32p]),
    "lmc_surrogate_set_params": (C.c_int, [_vp, _f32p]),
    "lmc_surrogate_set_normalization": (C.c_int, [_vp, _f64p, _f64p]),
    "lmc_surrogate_get_normalization": (C.c_int, [_vp, _f64p, _f64p]),
    "lmc_jump_samples": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "lmc_surrogate_forward": (C.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "lmc_surrogate_gradients": (C.c_int, [_vp, _vp, _vp, _i32, _vp, _f64p]),
    "lmc_surrogate_train": (C.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, C.c_double, _u64,
                                      _f64p]),
    "lmc_surrogate_eval_jumps": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, C.c_double,
                                           C.c_double, _vp, _vp, _vp, _vp, _vp]),
    "lmc_kmc_set_surrogate": (C.c_int, [_vp, _vp, C.c_double, C.c_double]),
    "lmc_kmc_surrogate_stats": (C.c_int, [_vp, _u64p, _u64p]),
    "lmc_kmc_run_graph": (C.c_int, [_vp, _u64, _i32, _vp]),
    "lmc_kmc_build_graph": (C.c_int, [_vp, _u64, _i32]),
}

_lib = None


class LmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[lmc status {code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m latticemontecarlo_amd.build` (hipcc, gfx950)")
        # torch ships its own libamdhip64.so.7; load it first so that liblmc.so binds to the
        # same HIP runtime instance as the torch tensors it receives.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != LMC_OK:
        raise LmcError(rc, lib().lmc_last_error().decode())
    return rc


def ptr(a, t):
    return a.ctypes.data_as(t)


def dptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
