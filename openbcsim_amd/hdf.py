"""The reference's HDF5 file formats for phantoms, beam profiles and scans, read
and written with the package's own HDF5 code (openbcsim_amd.h5), and their
transfer to a simulator.

Dataset names, shapes and checks follow src/utils/HDFConvenience.cpp:
  fixed scatterers   "data" (N, 4): x, y, z, amplitude                     (:45-66)
  spline scatterers  "control_points" (N, M, 3), "amplitudes" (N,),
                     "knot_vector" (M + p + 1,), "spline_degree" (scalar)  (:68-104)
  scan sequence      "origins", "directions", "lateral_dirs" (L, 3),
                     "timestamps" (L,), "lengths" (scalar line length)      (:106-128)
  beam profile       "beam_profile" (n_r, n_l, n_e), "rad_extent",
                     "lat_extent", "ele_extent" ([min, max] each); the
                     samples are divided by their maximum on loading        (:140-196)
  excitation         "samples", "center_index", "sampling_frequency"        (:130-138)
The reference's Python scripts write and read the same names with h5py
(e.g. phantom_scripts/cyst_phantom_2d.py:35-40, contracting_cylinder_spline.py:47-51).
Values are read as float32 (int for the spline degree and centre index),
whatever numeric type the file stores, as the reference's reader converts them.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import h5

__all__ = ["SplineScatterers", "ScanSequence", "BeamProfile", "Excitation",
           "load_fixed_scatterers", "load_spline_scatterers", "load_scan_sequence", "load_beam_profile",
           "load_excitation", "save_fixed_scatterers", "save_spline_scatterers", "save_scan_sequence",
           "save_beam_profile", "save_excitation", "set_fixed_scatterers_from_hdf",
           "add_spline_scatterers_from_hdf", "set_scan_sequence_from_hdf", "set_beam_profile_from_hdf"]


class SplineScatterers(NamedTuple):
    control_points: np.ndarray  # (N, M, 3) float32
    amplitudes: np.ndarray      # (N,) float32
    knot_vector: np.ndarray     # float32
    spline_degree: int


class ScanSequence(NamedTuple):
    origins: np.ndarray         # (L, 3) float32
    directions: np.ndarray      # (L, 3) float32
    lateral_dirs: np.ndarray    # (L, 3) float32
    timestamps: np.ndarray      # (L,) float32
    line_length: float


class BeamProfile(NamedTuple):
    samples: np.ndarray         # (n_r, n_l, n_e) float32, maximum 1
    rad_extent: tuple
    lat_extent: tuple
    ele_extent: tuple


class Excitation(NamedTuple):
    samples: np.ndarray         # float32
    center_index: int
    sampling_frequency: float


def _get(d: dict, name: str, path: str) -> np.ndarray:
    if name not in d:
        raise RuntimeError(f"{path}: no dataset '{name}'")
    return d[name]


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _scalar(a: np.ndarray, what: str):
    a = np.asarray(a)
    if a.size != 1:
        raise RuntimeError(f"'{what}' must hold one value")
    return a.reshape(()).item()


# --------------------------------------------------------------------------- loading
def load_fixed_scatterers(path: str) -> np.ndarray:
    """(N, 4) float32 rows x, y, z, amplitude (HDFConvenience.cpp:45-66)."""
    data = _get(h5.read(path), "data", path)
    if data.ndim != 2 or data.shape[1] != 4:
        raise RuntimeError("failed to load fixed scatterers: second dimension must have length four")
    return _f32(data)


def load_spline_scatterers(path: str) -> SplineScatterers:
    """Control points, amplitudes, knot vector and degree (HDFConvenience.cpp:68-104)."""
    d = h5.read(path)
    cp = _get(d, "control_points", path)
    if cp.ndim != 3 or cp.shape[2] != 3:
        raise RuntimeError("failed to load spline scatterers: illegal number of components (should be 3)")
    amps = _get(d, "amplitudes", path).reshape(-1)
    if amps.shape[0] != cp.shape[0]:
        raise RuntimeError("failed to load spline scatterers: amplitudes and control_points disagree in count")
    knots = _get(d, "knot_vector", path).reshape(-1)
    degree = int(_scalar(_get(d, "spline_degree", path), "spline_degree"))
    return SplineScatterers(_f32(cp), _f32(amps), _f32(knots), degree)


def load_scan_sequence(path: str) -> ScanSequence:
    """Scan lines and the common line length (HDFConvenience.cpp:106-128)."""
    d = h5.read(path)
    dirs, origins = _get(d, "directions", path), _get(d, "origins", path)
    lats, ts = _get(d, "lateral_dirs", path), _get(d, "timestamps", path).reshape(-1)
    length = float(np.float32(_scalar(_get(d, "lengths", path), "lengths")))
    n = dirs.shape[0]
    if origins.shape[0] != n or ts.shape[0] != n:
        raise RuntimeError("ScanSequence::size error")
    return ScanSequence(_f32(origins), _f32(dirs), _f32(lats), _f32(ts), length)


def load_beam_profile(path: str) -> BeamProfile:
    """LUT samples divided by their maximum, and the three extents
    (HDFConvenience.cpp:140-196)."""
    d = h5.read(path)
    s = _get(d, "beam_profile", path)
    if s.ndim != 3:
        raise RuntimeError("beam_profile data must be a 3D array")
    ext = []
    for name in ("rad_extent", "lat_extent", "ele_extent"):
        e = _f32(_get(d, name, path)).reshape(-1)
        if e.shape[0] != 2:
            raise RuntimeError(f"{name} must contain two elements: [min, max]")
        ext.append((float(e[0]), float(e[1])))
    s = _f32(s)
    return BeamProfile(s / s.max(), *ext)


def load_excitation(path: str) -> Excitation:
    """Excitation samples, centre index and sampling frequency
    (HDFConvenience.cpp:130-138 reads the same three names; the demodulation
    frequency is not stored in the file)."""
    d = h5.read(path)
    return Excitation(_f32(_get(d, "samples", path)).reshape(-1),
                      int(_scalar(_get(d, "center_index", path), "center_index")),
                      float(np.float32(_scalar(_get(d, "sampling_frequency", path), "sampling_frequency"))))


# --------------------------------------------------------------------------- saving
def save_fixed_scatterers(path: str, data) -> None:
    a = _f32(data)
    if a.ndim != 2 or a.shape[1] != 4:
        raise RuntimeError("fixed scatterers must be an (N, 4) array")
    h5.write(path, {"data": a})


def save_spline_scatterers(path: str, control_points, amplitudes, knot_vector, spline_degree: int) -> None:
    h5.write(path, {"spline_degree": np.int64(spline_degree), "control_points": _f32(control_points),
                    "amplitudes": _f32(amplitudes), "knot_vector": _f32(knot_vector)})


def save_scan_sequence(path: str, origins, directions, lateral_dirs, timestamps, line_length: float) -> None:
    h5.write(path, {"origins": _f32(origins), "directions": _f32(directions), "lateral_dirs": _f32(lateral_dirs),
                    "timestamps": _f32(timestamps), "lengths": np.float32(line_length)})


def save_beam_profile(path: str, samples, rad_extent, lat_extent, ele_extent) -> None:
    h5.write(path, {"beam_profile": _f32(samples), "rad_extent": _f32(rad_extent),
                    "lat_extent": _f32(lat_extent), "ele_extent": _f32(ele_extent)})


def save_excitation(path: str, samples, center_index: int, sampling_frequency: float) -> None:
    h5.write(path, {"samples": _f32(samples), "center_index": np.int64(center_index),
                    "sampling_frequency": np.float32(sampling_frequency)})


# --------------------------------------------------------------------------- to a simulator
def set_fixed_scatterers_from_hdf(sim, path: str) -> None:
    """Replace the simulator's fixed scatterers with the file's."""
    sim.clear_fixed_scatterers()
    sim.add_fixed_scatterers(load_fixed_scatterers(path))


def add_spline_scatterers_from_hdf(sim, path: str) -> None:
    s = load_spline_scatterers(path)
    sim.add_spline_scatterers(s.spline_degree, s.knot_vector, s.control_points, s.amplitudes)


def set_scan_sequence_from_hdf(sim, path: str) -> None:
    s = load_scan_sequence(path)
    sim.set_scan_sequence(s.origins, s.directions, s.line_length, s.lateral_dirs, s.timestamps)


def set_beam_profile_from_hdf(sim, path: str) -> None:
    b = load_beam_profile(path)
    sim.set_lut_beam_profile(b.rad_extent[0], b.rad_extent[1], b.lat_extent[0], b.lat_extent[1],
                             b.ele_extent[0], b.ele_extent[1], b.samples)
