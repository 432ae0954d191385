"""Python mirror of the reference's ``pyrfsim.RfSimulator`` over the C-ABI.

Method names, argument order, shapes and the ``(num_samples, num_lines)``
complex64 result of ``simulate_lines`` follow the reference binding
(src/python/PythonInterface.cpp:58-351), so the reference's Python scripts can
drive the MI355X path by importing ``RfSimulator`` from here.  Errors raise
``RuntimeError`` subclasses, as the Boost.Python binding turns the reference's
``std::runtime_error`` into ``RuntimeError``.

Additions beyond the reference binding: ``simulate_frame`` (``[line][sample]``
layout, no transpose), ``simulate_lines_device`` (IQ left in device memory),
``get_projection`` (time-projection bins) and ``close``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BcsimError, check

__all__ = ["RfSimulator", "BcsimError"]


def _f32(a, ndim=None, name="array"):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if ndim is not None and a.ndim != ndim:
        raise RuntimeError(f"{name}: invalid rank {a.ndim}, expected {ndim}")
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class RfSimulator:
    """The "gpu" simulator (PythonInterface.cpp:58-323)."""

    def __init__(self, sim_type: str = "gpu"):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.bcsim_create(sim_type.encode(), C.byref(h)))
        self._h = h
        self._log_cb = None
        self._print_debug = False

    # --- lifetime -----------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.bcsim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # --- IAlgorithm -----------------------------------------------------------------------
    def set_print_debug(self, val: bool):
        self._print_debug = bool(val)

    def set_parameter(self, key: str, value: str):
        check(self._lib.bcsim_set_parameter(self._h, str(key).encode(), str(value).encode()))

    def get_parameter(self, key: str) -> str:
        size = 512
        while True:
            buf = C.create_string_buffer(size)
            status = self._lib.bcsim_get_parameter(self._h, str(key).encode(), buf, len(buf))
            if status == 5 and size < (1 << 24):  # BCSIM_ERR_CAPACITY: long values (launch plans, tuning reports)
                size *= 8
                continue
            check(status)
            return buf.value.decode()

    def clear_fixed_scatterers(self):
        check(self._lib.bcsim_clear_fixed_scatterers(self._h))

    def add_fixed_scatterers(self, data):
        a, p = _f32(data, 2, "add_fixed_scatterers")
        if a.shape[1] != 4:
            raise RuntimeError("add_fixed_scatterers : Number of columns must be four")
        check(self._lib.bcsim_add_fixed_scatterers(self._h, p, a.shape[0]))

    def set_fixed_scatterers(self, data):
        a, p = _f32(data, 2, "set_fixed_scatterers")
        if a.shape[1] != 4:
            raise RuntimeError("set_fixed_scatterers : Number of columns must be four")
        check(self._lib.bcsim_set_fixed_scatterers(self._h, p, a.shape[0]))

    def clear_spline_scatterers(self):
        check(self._lib.bcsim_clear_spline_scatterers(self._h))

    def add_spline_scatterers(self, spline_degree, knot_vector, control_points, amplitudes):
        k, pk = _f32(knot_vector, 1, "knot vector")
        c, pc = _f32(control_points, 3, "control_points")
        a, pa = _f32(amplitudes, 1, "amplitudes")
        if c.shape[2] != 3:
            raise RuntimeError("add_spline_scatterers : size of dimension 3 must be three (x,y,z)")
        if a.shape[0] != c.shape[0]:
            raise RuntimeError("Mismatch between control_points and amplitudes")
        check(self._lib.bcsim_add_spline_scatterers(self._h, int(spline_degree), pk, k.shape[0], pc, pa,
                                                    c.shape[0], c.shape[1]))

    def set_scan_sequence(self, origins, directions, line_length, lateral_dirs, timestamps,
                          all_timestamps_equal=False):
        o, po = _f32(origins, 2, "origins")
        d, pd = _f32(directions, 2, "directions")
        l, pl = _f32(lateral_dirs, 2, "lateral_dirs")
        t, pt = _f32(timestamps, 1, "timestamps")
        n = o.shape[0]
        if d.shape[0] != n or l.shape[0] != n or t.shape[0] != n:
            raise RuntimeError("set_scan_sequence(): Mismatch in number of rows.")
        check(self._lib.bcsim_set_scan_sequence(self._h, n, po, pd, pl, pt, float(line_length),
                                                1 if all_timestamps_equal else 0))

    def set_excitation(self, samples, center_index, fs, demod_freq):
        s, p = _f32(samples, 1, "set_excitation: samples")
        check(self._lib.bcsim_set_excitation(self._h, p, s.shape[0], int(center_index), float(fs),
                                             float(demod_freq)))

    def set_analytical_beam_profile(self, sigma_lateral, sigma_elevational):
        check(self._lib.bcsim_set_analytical_profile(self._h, float(sigma_lateral), float(sigma_elevational)))

    def set_lut_beam_profile(self, r_min, r_max, l_min, l_max, e_min, e_max, samples):
        s, p = _f32(samples, None, "samples")
        if s.ndim != 3:
            raise RuntimeError("Number of sample dimensions must be three.")
        check(self._lib.bcsim_set_lut_profile(self._h, float(r_min), float(r_max), float(l_min), float(l_max),
                                              float(e_min), float(e_max), p, s.shape[0], s.shape[1], s.shape[2]))

    def output_shape(self):
        L, S = C.c_int(), C.c_int()
        check(self._lib.bcsim_get_output_shape(self._h, C.byref(L), C.byref(S)))
        return L.value, S.value

    def simulate_frame(self, out: np.ndarray = None) -> np.ndarray:
        """IQ frame as complex64 [line][sample] (into `out` when given: C-contiguous
        complex64 of at least lines*samples elements, returned reshaped)."""
        L, S = self.output_shape()
        if out is None:
            out = np.empty((L, S), np.complex64)
        else:
            if out.dtype != np.complex64 or not out.flags.c_contiguous or out.size < L * S:
                raise ValueError("out must be a C-contiguous complex64 array of at least lines*samples elements")
            out = out.reshape(-1)[: L * S].reshape(L, S)
        nl, ns = C.c_int(), C.c_int()
        check(self._lib.bcsim_simulate_lines(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                             C.byref(nl), C.byref(ns)))
        return out

    def simulate_lines(self) -> np.ndarray:
        """IQ frame as complex64 (num_samples, num_lines), like the reference binding."""
        return np.ascontiguousarray(self.simulate_frame().T)

    def simulate_lines_device(self, device_ptr: int, capacity: int):
        """Enqueue a frame whose IQ goes to device memory at device_ptr; returns (L, S)."""
        nl, ns = C.c_int(), C.c_int()
        check(self._lib.bcsim_simulate_lines_device(self._h, C.c_void_p(device_ptr), capacity,
                                                    C.byref(nl), C.byref(ns)))
        return nl.value, ns.value

    def synchronize(self):
        check(self._lib.bcsim_synchronize(self._h))

    def get_projection(self) -> np.ndarray:
        """Time-projection bins of the last frame, complex64 [line][num_rf]."""
        nl, nr = C.c_int(), C.c_int()
        check(self._lib.bcsim_debug_get_projection(self._h, None, 0, C.byref(nl), C.byref(nr)))
        out = np.empty((nl.value, nr.value), np.complex64)
        check(self._lib.bcsim_debug_get_projection(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                                   C.byref(nl), C.byref(nr)))
        return out

    def get_debug_data(self, identifier: str):
        n = C.c_size_t()
        check(self._lib.bcsim_get_debug_data(self._h, identifier.encode(), None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.float64)
        check(self._lib.bcsim_get_debug_data(self._h, identifier.encode(),
                                             out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return list(out[: n.value])

    def get_total_num_scatterers(self) -> int:
        n = C.c_size_t()
        check(self._lib.bcsim_get_total_num_scatterers(self._h, C.byref(n)))
        return n.value

    def set_logger(self, fn):
        """fn(level:int, msg:str) or None."""
        if fn is None:
            self._log_cb = None
            check(self._lib.bcsim_set_logger(self._h, _lib.LOG_FN(), None))
            return
        self._log_cb = _lib.LOG_FN(lambda lvl, msg, user: fn(lvl, msg.decode()))
        check(self._lib.bcsim_set_logger(self._h, self._log_cb, None))
