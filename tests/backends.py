"""ctypes drivers for the two checkers: the FP64 oracle (oracle/liboracle.so) and
the compiled reference CPU algorithm (oracle/_ref/libbcsim_ref.so).

Both expose the same method names as ``openbcsim_amd.pyrfsim.RfSimulator`` so a
``Workload`` configures all three identically.  ``simulate_frame`` returns
``[line][sample]`` complex arrays (complex128 for the oracle, complex64 for the
reference); ``project`` returns the time-projection bins ``[line][num_rf]``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libbcsim_ref.so")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_create.restype = C.c_void_p
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_last_error.argtypes = [C.c_void_p]
        lib.oracle_destroy.argtypes = [C.c_void_p]
        for name in ["oracle_set_parameter", "oracle_set_excitation", "oracle_set_scan_sequence",
                     "oracle_add_fixed_scatterers", "oracle_add_spline_scatterers",
                     "oracle_set_analytical_profile", "oracle_set_lut_profile",
                     "oracle_get_output_shape", "oracle_simulate_lines", "oracle_project",
                     "oracle_clear_fixed_scatterers", "oracle_clear_spline_scatterers"]:
            getattr(lib, name).restype = C.c_int
        lib.oracle_set_parameter.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        lib.oracle_set_excitation.argtypes = [C.c_void_p, _fp, C.c_int, C.c_int, C.c_float, C.c_float]
        lib.oracle_set_scan_sequence.argtypes = [C.c_void_p, C.c_int, _fp, _fp, _fp, _fp, C.c_float]
        lib.oracle_add_fixed_scatterers.argtypes = [C.c_void_p, _fp, C.c_size_t]
        lib.oracle_add_spline_scatterers.argtypes = [C.c_void_p, C.c_int, _fp, C.c_int, _fp, _fp,
                                                     C.c_size_t, C.c_int]
        lib.oracle_clear_fixed_scatterers.argtypes = [C.c_void_p]
        lib.oracle_clear_spline_scatterers.argtypes = [C.c_void_p]
        lib.oracle_set_analytical_profile.argtypes = [C.c_void_p, C.c_float, C.c_float]
        lib.oracle_set_lut_profile.argtypes = [C.c_void_p] + [C.c_float] * 6 + [_fp, C.c_int, C.c_int, C.c_int]
        ip = C.POINTER(C.c_int)
        lib.oracle_get_output_shape.argtypes = [C.c_void_p, ip, ip, ip, ip]
        lib.oracle_simulate_lines.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_int]
        lib.oracle_project.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_int]
        lib.oracle_next_pow2.restype = C.c_size_t
        lib.oracle_next_pow2.argtypes = [C.c_size_t]
        lib.oracle_hilbert_mask.argtypes = [C.c_int, _dp]
        lib.oracle_fft.argtypes = [_dp, C.c_int, C.c_int]
        lib.oracle_fft_conv.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp]
        lib.oracle_num_rf_samples.restype = C.c_int
        lib.oracle_num_rf_samples.argtypes = [C.c_float, C.c_float, C.c_float]
        lib.oracle_bspline_basis.restype = C.c_double
        lib.oracle_bspline_basis.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_int]
        lib.oracle_bspline_basis_f32.restype = C.c_float
        lib.oracle_bspline_basis_f32.argtypes = [C.c_int, C.c_int, C.c_float, _fp, C.c_int]
        lib.oracle_knot_interval.restype = C.c_int
        lib.oracle_knot_interval.argtypes = [_fp, C.c_int, C.c_float]
        lib.oracle_contract_bin.restype = C.c_int
        lib.oracle_contract_bin.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_float, C.c_float, C.c_int, _fp]
        lib.oracle_lut_sample.restype = C.c_double
        lib.oracle_lut_sample.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_create.restype = C.c_void_p
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_last_error.argtypes = [C.c_void_p]
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_set_parameter.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        lib.ref_set_excitation.argtypes = [C.c_void_p, _fp, C.c_int, C.c_int, C.c_float, C.c_float]
        lib.ref_set_scan_sequence.argtypes = [C.c_void_p, C.c_int, _fp, _fp, _fp, _fp, C.c_float]
        lib.ref_add_fixed_scatterers.argtypes = [C.c_void_p, _fp, C.c_size_t]
        lib.ref_add_spline_scatterers.argtypes = [C.c_void_p, C.c_int, _fp, C.c_int, _fp, _fp,
                                                  C.c_size_t, C.c_int]
        lib.ref_clear_fixed_scatterers.argtypes = [C.c_void_p]
        lib.ref_clear_spline_scatterers.argtypes = [C.c_void_p]
        lib.ref_set_analytical_profile.argtypes = [C.c_void_p, C.c_float, C.c_float]
        lib.ref_set_lut_profile.argtypes = [C.c_void_p] + [C.c_float] * 6 + [_fp, C.c_int, C.c_int, C.c_int]
        ip = C.POINTER(C.c_int)
        lib.ref_get_output_shape.argtypes = [C.c_void_p, ip, ip, ip]
        lib.ref_simulate_lines.argtypes = [C.c_void_p, _fp, C.c_size_t, ip, ip]
        lib.ref_project.argtypes = [C.c_void_p, _fp, C.c_size_t]
        _ref = lib
    return _ref


class _Base:
    prefix = ""

    def _call(self, name, *args):
        rc = getattr(self.lib, self.prefix + name)(self.h, *args)
        if rc != 0:
            raise RuntimeError(getattr(self.lib, self.prefix + "last_error")(self.h).decode())

    def set_parameter(self, key, value):
        self._call("set_parameter", str(key).encode(), str(value).encode())

    def set_excitation(self, samples, center_index, fs, demod_freq):
        a, p = _f32(samples)
        self._call("set_excitation", p, len(a), int(center_index), float(fs), float(demod_freq))

    def set_scan_sequence(self, origins, directions, line_length, lateral_dirs, timestamps,
                          all_timestamps_equal=False):
        o, po = _f32(origins)
        d, pd = _f32(directions)
        l, pl = _f32(lateral_dirs)
        t, pt = _f32(timestamps)
        self._call("set_scan_sequence", o.shape[0], po, pd, pl, pt, float(line_length))

    def add_fixed_scatterers(self, data):
        a, p = _f32(data)
        self._call("add_fixed_scatterers", p, a.shape[0])

    def clear_fixed_scatterers(self):
        self._call("clear_fixed_scatterers")

    def clear_spline_scatterers(self):
        self._call("clear_spline_scatterers")

    def add_spline_scatterers(self, degree, knots, control_points, amplitudes):
        k, pk = _f32(knots)
        c, pc = _f32(control_points)
        a, pa = _f32(amplitudes)
        self._call("add_spline_scatterers", int(degree), pk, len(k), pc, pa, c.shape[0], c.shape[1])

    def set_analytical_beam_profile(self, sigma_lateral, sigma_elevational):
        self._call("set_analytical_profile", float(sigma_lateral), float(sigma_elevational))

    def set_lut_beam_profile(self, r0, r1, l0, l1, e0, e1, samples):
        s, ps = _f32(samples)
        self._call("set_lut_profile", float(r0), float(r1), float(l0), float(l1), float(e0),
                   float(e1), ps, s.shape[0], s.shape[1], s.shape[2])

    def close(self):
        if self.h:
            getattr(self.lib, self.prefix + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Oracle(_Base):
    """FP64 COLE oracle."""
    prefix = "oracle_"

    def __init__(self, num_threads: int = 0):
        self.lib = oracle_lib()
        self.h = C.c_void_p(self.lib.oracle_create())
        self.num_threads = num_threads

    def shape(self):
        L, S, R, N = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._call("get_output_shape", C.byref(L), C.byref(S), C.byref(R), C.byref(N))
        return L.value, S.value, R.value, N.value

    def simulate_frame(self) -> np.ndarray:
        L, S, _, _ = self.shape()
        out = np.zeros((L, S), np.complex128)
        self._call("simulate_lines", out.ctypes.data_as(_dp), out.size, self.num_threads)
        return out

    def project(self) -> np.ndarray:
        L, _, R, _ = self.shape()
        out = np.zeros((L, R), np.complex128)
        self._call("project", out.ctypes.data_as(_dp), out.size, self.num_threads)
        return out

    def lut_sample(self, r, l, e) -> float:
        return self.lib.oracle_lut_sample(self.h, float(r), float(l), float(e))


class Reference(_Base):
    """The reference's CpuAlgorithm, compiled from its own sources."""
    prefix = "ref_"

    def __init__(self):
        self.lib = ref_lib()
        self.h = C.c_void_p(self.lib.ref_create())

    def shape(self):
        L, S, R = C.c_int(), C.c_int(), C.c_int()
        self._call("get_output_shape", C.byref(L), C.byref(S), C.byref(R))
        return L.value, S.value, R.value

    def simulate_frame(self) -> np.ndarray:
        L, S, _ = self.shape()
        out = np.zeros((L, S), np.complex64)
        nl, ns = C.c_int(), C.c_int()
        self._call("simulate_lines", out.ctypes.data_as(_fp), out.size, C.byref(nl), C.byref(ns))
        assert (nl.value, ns.value) == (L, S), ((nl.value, ns.value), (L, S))
        return out

    def project(self) -> np.ndarray:
        L, _, R = self.shape()
        out = np.zeros((L, R), np.complex64)
        self._call("project", out.ctypes.data_as(_fp), out.size)
        return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / den) if den > 0 else float(np.linalg.norm(a.ravel()))
