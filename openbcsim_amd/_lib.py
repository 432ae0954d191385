"""Loader for libbcsim_hip.so: the ctypes signatures of include/bcsim_c.h.

The library is built in-tree by openbcsim_amd.build; loading fails loudly if it
is missing -- there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbcsim_hip.so")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_sp = C.POINTER(C.c_size_t)
_h = C.c_void_p

LOG_FN = C.CFUNCTYPE(None, C.c_int, C.c_char_p, C.c_void_p)

# name -> argtypes (every function returns int status except bcsim_last_error)
SIGNATURES = {
    "bcsim_create": [C.c_char_p, C.POINTER(C.c_void_p)],
    "bcsim_create_simulator": [C.c_char_p, C.POINTER(C.c_void_p)],
    "bcsim_destroy": [_h],
    "bcsim_set_parameter": [_h, C.c_char_p, C.c_char_p],
    "bcsim_get_parameter": [_h, C.c_char_p, C.c_char_p, C.c_size_t],
    "bcsim_set_excitation": [_h, _fp, C.c_int, C.c_int, C.c_float, C.c_float],
    "bcsim_set_scan_sequence": [_h, C.c_int, _fp, _fp, _fp, _fp, C.c_float, C.c_int],
    "bcsim_clear_fixed_scatterers": [_h],
    "bcsim_add_fixed_scatterers": [_h, _fp, C.c_size_t],
    "bcsim_set_fixed_scatterers": [_h, _fp, C.c_size_t],
    "bcsim_clear_spline_scatterers": [_h],
    "bcsim_add_spline_scatterers": [_h, C.c_int, _fp, C.c_int, _fp, _fp, C.c_size_t, C.c_int],
    "bcsim_set_spline_scatterers": [_h, C.c_int, _fp, C.c_int, _fp, _fp, C.c_size_t, C.c_int],
    "bcsim_set_analytical_profile": [_h, C.c_float, C.c_float],
    "bcsim_set_lut_profile": [_h] + [C.c_float] * 6 + [_fp, C.c_int, C.c_int, C.c_int],
    "bcsim_get_output_shape": [_h, _ip, _ip],
    "bcsim_simulate_lines": [_h, _fp, C.c_size_t, _ip, _ip],
    "bcsim_simulate_lines_device": [_h, C.c_void_p, C.c_size_t, _ip, _ip],
    "bcsim_simulate_lines_device_noalloc": [_h, C.c_void_p, C.c_size_t, _ip, _ip],
    "bcsim_synchronize": [_h],
    "bcsim_get_stream": [_h, C.POINTER(C.c_void_p)],
    "bcsim_get_debug_data": [_h, C.c_char_p, _dp, C.c_size_t, _sp],
    "bcsim_get_total_num_scatterers": [_h, _sp],
    "bcsim_set_logger": [_h, LOG_FN, C.c_void_p],
    "bcsim_debug_get_projection": [_h, _fp, C.c_size_t, _ip, _ip],
    "bcsim_batch_create": [C.POINTER(C.c_void_p)],
    "bcsim_batch_destroy": [_h],
    "bcsim_batch_set_parameter": [_h, C.c_char_p, C.c_char_p],
    "bcsim_batch_add": [_h, _h, _ip],
    "bcsim_batch_clear": [_h],
    "bcsim_batch_get_layout": [_h, _sp, _ip, _sp, _ip, _ip, C.c_int],
    "bcsim_batch_simulate_device": [_h, C.c_void_p, C.c_size_t],
    "bcsim_batch_simulate": [_h, _fp, C.c_size_t],
    "bcsim_batch_synchronize": [_h],
    "bcsim_batch_get_debug_data": [_h, C.c_char_p, _dp, C.c_size_t, _sp],
    "bcsim_server_create": [C.POINTER(C.c_void_p)],
    "bcsim_server_destroy": [_h],
    "bcsim_server_set_parameter": [_h, C.c_char_p, C.c_char_p],
    "bcsim_server_add_phantom": [_h, _h, _ip],
    "bcsim_server_add_geometry": [_h, C.c_int, _fp, _fp, _fp, C.c_float, _ip],
    "bcsim_server_get_output_shape": [_h, C.c_int, C.c_int, _ip, _ip],
    "bcsim_server_submit": [_h, C.c_int, C.c_int, C.c_float, _fp, C.c_size_t, C.POINTER(C.c_uint64)],
    "bcsim_server_wait": [_h, C.c_uint64, C.c_int, _ip, _ip, _dp],
    "bcsim_server_get_stats": [_h, C.c_char_p, _dp, C.c_size_t, _sp],
    "bcsim_server_alloc_host": [_h, C.c_size_t, C.POINTER(C.c_void_p)],
    "bcsim_server_free_host": [_h, C.c_void_p],
}

_lib = None


def load(path: str = LIB_PATH):
    """Load and type the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python -m openbcsim_amd.build` "
                           "(the simulation path has no CPU fallback)")
    lib = C.CDLL(path)
    lib.bcsim_last_error.restype = C.c_char_p
    lib.bcsim_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    _lib = lib
    return lib


class BcsimError(RuntimeError):
    """A failed C-ABI call; .status is the bcsim_status code."""

    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


def check(status: int) -> None:
    if status != 0:
        raise BcsimError(status, load().bcsim_last_error().decode())
