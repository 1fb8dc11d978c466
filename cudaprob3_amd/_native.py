"""Loader for the native libraries built in-tree by ``cudaprob3_amd/csrc``.

``libcp3hip.so`` is the product (HIP kernels behind the C ABI declared in
``include/cp3_hip.h``).  There is no fallback: if the library is missing the
import of any propagator raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_SO = os.path.join(PKG, "libcp3hip.so")
# CP3_MODEL_SO selects another build of the host model library (the
# sanitizer build, libcp3model_asan.so, for the CPU tests)
MODEL_SO = os.environ.get("CP3_MODEL_SO") or os.path.join(PKG, "libcp3model.so")

CP3_FP64, CP3_FP32, CP3_MIXED = 0, 1, 2
CP3_SOLVE_AUTO, CP3_SOLVE_SHARED, CP3_SOLVE_TABLE = 0, 1, 2
SOLVE_MODES = {"auto": CP3_SOLVE_AUTO, "shared": CP3_SOLVE_SHARED,
               "table": CP3_SOLVE_TABLE}
PRECISIONS = {"fp64": CP3_FP64, "fp32": CP3_FP32, "mixed": CP3_MIXED}

CP3_OK = 0
CP3_ERR_INVALID = -1
CP3_ERR_STATE = -2
CP3_ERR_RANGE = -3
CP3_ERR_DEVICE = -4
CP3_ERR_NOMEM = -5
CP3_ERR_CAPACITY = -6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class Cp3Error(RuntimeError):
    """A failing C ABI call: ``code`` is the CP3_ERR_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


# name -> (restype, argtypes); every function returns int unless noted.
_SIGS = {
    "cp3_last_error": (C.c_char_p, []),
    "cp3_version": (C.c_int, []),
    "cp3_device_count": (C.c_int, [_ip]),
    "cp3_atmos_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "cp3_atmos_destroy": (C.c_int, [_vp]),
    "cp3_atmos_set_energy_list": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_atmos_set_cosine_list": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_atmos_set_mns_matrix": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "cp3_atmos_set_neutrino_masses": (C.c_int, [_vp, C.c_double, C.c_double]),
    "cp3_atmos_set_density": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int]),
    "cp3_atmos_set_density_from_file": (C.c_int, [_vp, C.c_char_p]),
    "cp3_atmos_set_chemical_composition": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_atmos_set_production_height": (C.c_int, [_vp, C.c_double]),
    "cp3_atmos_get_nlayer_boundaries": (C.c_int, [_vp, _ip]),
    "cp3_atmos_calculate": (C.c_int, [_vp, C.c_int]),
    "cp3_atmos_calculate_async": (C.c_int, [_vp, C.c_int]),
    "cp3_atmos_calculate_async_static": (C.c_int, [_vp, C.c_int, _vp]),
    "cp3_atmos_wait": (C.c_int, [_vp]),
    "cp3_atmos_get_probability": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "cp3_atmos_get_probability_arr": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_atmos_result_device": (C.c_int, [_vp, C.POINTER(_vp), _ip]),
    "cp3_atmos_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "cp3_atmos_last_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "cp3_atmos_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cp3_atmos_set_solve_mode": (C.c_int, [_vp, C.c_int]),
    "cp3_atmos_set_density_poly": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    "cp3_atmos_set_production_height_bins": (C.c_int, [_vp, C.c_int]),
    "cp3_atmos_set_production_height_averaging": (C.c_int, [_vp, C.c_int]),
    "cp3_atmos_set_production_height_list": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int]),
    "cp3_beam_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "cp3_beam_destroy": (C.c_int, [_vp]),
    "cp3_beam_set_energy_list": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_beam_set_mns_matrix": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "cp3_beam_set_neutrino_masses": (C.c_int, [_vp, C.c_double, C.c_double]),
    "cp3_beam_set_density": (C.c_int, [_vp, C.c_double]),
    "cp3_beam_set_path_length": (C.c_int, [_vp, C.c_double]),
    "cp3_beam_calculate": (C.c_int, [_vp, C.c_int]),
    "cp3_beam_calculate_async": (C.c_int, [_vp, C.c_int]),
    "cp3_beam_wait": (C.c_int, [_vp]),
    "cp3_beam_get_probability": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "cp3_beam_get_probability_arr": (C.c_int, [_vp, _dp, C.c_int]),
    "cp3_beam_result_device": (C.c_int, [_vp, C.POINTER(_vp), _ip]),
    "cp3_beam_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "cp3_beam_last_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
}


class Cp3Job(C.Structure):
    """struct cp3_job (include/cp3_hip.h)."""
    _fields_ = [("kind", C.c_int32), ("grid", C.c_int32),
                ("nutype", C.c_int32), ("knutype", C.c_int32),
                ("theta12", C.c_double), ("theta13", C.c_double),
                ("theta23", C.c_double), ("dcp", C.c_double),
                ("dm12sq", C.c_double), ("dm23sq", C.c_double)]


_i64p = C.POINTER(C.c_int64)
_SIGS.update({
    "cp3_batch_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "cp3_batch_destroy": (C.c_int, [_vp]),
    "cp3_batch_add_atmos_grid": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _ip]),
    "cp3_batch_add_beam_grid": (C.c_int, [_vp, _dp, C.c_int, C.c_double, C.c_double, _ip]),
    "cp3_batch_set_jobs": (C.c_int, [_vp, C.POINTER(Cp3Job), C.c_int]),
    "cp3_batch_update_jobs": (C.c_int, [_vp, C.POINTER(Cp3Job), C.c_int]),
    "cp3_batch_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "cp3_batch_autotune": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip]),
    "cp3_batch_set_solve_mode": (C.c_int, [_vp, C.c_int]),
    "cp3_batch_phase_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "cp3_batch_get_tuning": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "cp3_batch_use_graph": (C.c_int, [_vp, C.c_int]),
    "cp3_batch_run": (C.c_int, [_vp]),
    "cp3_batch_wait": (C.c_int, [_vp]),
    "cp3_batch_last_ms": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "cp3_batch_result_device": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), _ip]),
    "cp3_batch_copy_result": (C.c_int, [_vp, C.c_int, _dp]),
    "cp3_batch_stats": (C.c_int, [_vp, _i64p, _i64p, _i64p]),
    "cp3_batch_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
})

_SIGS.update({
    "cp3_server_create": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.c_int, C.c_int, C.POINTER(_vp)]),
    "cp3_server_destroy": (C.c_int, [_vp]),
    "cp3_server_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cp3_server_set_solve_mode": (C.c_int, [_vp, C.c_int]),
    "cp3_server_add_atmos_grid": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _ip]),
    "cp3_server_add_beam_grid": (C.c_int, [_vp, _dp, C.c_int, C.c_double, C.c_double, _ip]),
    "cp3_server_submit": (C.c_int, [_vp, C.POINTER(Cp3Job), _i64p]),
    "cp3_server_wait": (C.c_int, [_vp, C.c_int64, C.c_int]),
    "cp3_server_fetch": (C.c_int, [_vp, C.c_int64, _dp]),
    "cp3_server_result_device": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp), _ip]),
    "cp3_server_release": (C.c_int, [_vp, C.c_int64]),
    "cp3_server_latency": (C.c_int, [_vp, C.c_int64, _dp, _dp]),
    "cp3_server_stats": (C.c_int, [_vp, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p]),
})

_lib = None


def load(path: str = HIP_SO):
    """Loads libcp3hip.so (once) and declares every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built: run `make -C cudaprob3_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols():
    return list(_SIGS)


def check(rc: int):
    if rc != CP3_OK:
        msg = load().cp3_last_error().decode(errors="replace")
        raise Cp3Error(rc, msg)
