"""Writes the golden fixtures in this directory from the reference's own CPU
algorithm (CpuAlgorithm, src/core/algorithm/CpuAlgorithm.cpp), compiled from the
reference sources by ``make -C oracle ref`` into oracle/_ref/libbcsim_ref.so.

Run where the reference tree exists:  python tests/golden/make_golden.py [case ...]

Each case is a reduced, seeded workload from openbcsim_amd.workloads; the
fixture stores the reference's IQ frame ([line][sample] complex64), its
time-projection bins in sparse form (flat index of every non-zero bin and its
complex64 value) and a checksum of the inputs so a drift in the generators is
caught instead of silently compared against stale outputs.
"""
from __future__ import annotations

import hashlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from openbcsim_amd import workloads as W  # noqa: E402


def golden_cases():
    """name -> Workload.  Small enough that the fixtures stay a few hundred KB."""
    cases = {}
    cases["fixed_gauss_arc"] = W.fixed_sector_2d(4000, 8, seed=11, line_length=0.04)
    cases["fixed_gauss_noarc_phase_dec3"] = W.fixed_sector_2d(4000, 8, seed=12, line_length=0.04,
                                                               phase_delay=True, arc=False, decimation=3)
    cases["spline_cubic_dec4"] = W.spline_mmode(3000, 8, seed=13, line_length=0.04, decimation=4)
    cases["spline_quadratic_phase"] = W.spline_mmode(2000, 6, seed=14, num_cp=7, degree=2,
                                                      line_length=0.04, phase_delay=True, decimation=2)
    wl = W.volume_lut(6000, n_az=3, n_el=3, seed=15, line_length=0.04, lut_shape=(32, 16, 16))
    cases["lut_volume_dec2"] = wl
    wl2 = W.volume_lut(6000, n_az=3, n_el=2, seed=16, line_length=0.04, lut_shape=(32, 16, 16),
                       phase_delay=True, decimation=1)
    cases["lut_volume_phase"] = wl2
    # two fixed datasets + one spline dataset on a linear scan (superposition across datasets)
    mix = W.fixed_sector_2d(1500, 6, seed=17, line_length=0.04, phase_delay=True)
    o, d, l, t, length = W.linear_scan(6, 0.02, 0.04)
    spl = W.spline_mmode(1500, 6, seed=18, line_length=0.04)
    t = spl.scan[3]
    mix.scan = (o, d, l, t, length)
    mix.fixed.append(W.fixed_sector_2d(1000, 6, seed=19).fixed[0])
    mix.splines = spl.splines
    mix.name = "mixed_linear_phase"
    cases["mixed_linear_phase"] = mix
    cases.update(full_size_cases())
    return cases


def _lines(wl, idx):
    o, d, l, t, length = wl.scan
    idx = np.asarray(idx)
    wl.scan = (o[idx], d[idx], l[idx], t[idx], length)
    return wl


def full_size_cases():
    """The benchmark configurations at their real sizes (phantom, line length, FFT
    length, decimation), restricted to four lines of their scans so the fixtures
    stay small: the 2D sector (100k points, 0.12 m, N_fft 8192, D=1) with phase
    delay off and on, the spline M-mode (200k cubic splines, 0.12 m, D=4) and
    the volume (1M points, 256x64x64 LUT, 0.15 m, N_fft 16384, D=2)."""
    return {
        "full_sector": _lines(W.fixed_sector_2d(), [10, 47, 64, 101]),
        "full_sector_phase": _lines(W.fixed_sector_2d(phase_delay=True), [5, 40, 63, 120]),
        "full_spline_mmode": _lines(W.spline_mmode(), [3, 77, 128, 250]),
        "full_volume": _lines(W.volume_lut(), [0, 1111, 2080, 4095]),
    }


def input_digest(wl) -> str:
    h = hashlib.sha256()
    for a in [wl.excitation, *wl.scan[:4]] + list(wl.fixed):
        h.update(np.ascontiguousarray(a).tobytes())
    for s in wl.splines:
        for a in (s.knots, s.control_points, s.amplitudes):
            h.update(np.ascontiguousarray(a).tobytes())
    if wl.profile[0] == "lut":
        h.update(np.ascontiguousarray(wl.profile[2]).tobytes())
    h.update(repr(sorted(wl.params.items())).encode())
    h.update(repr((wl.center_index, wl.fs, wl.demod_freq, float(wl.scan[4]))).encode())
    return h.hexdigest()


def main():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from backends import Reference  # noqa: E402

    only = set(sys.argv[1:])
    for name, wl in golden_cases().items():
        if only and name not in only:
            continue
        ref = Reference()
        wl.apply(ref)
        iq = ref.simulate_frame()
        bins = ref.project()
        flat = bins.ravel()
        nz = np.flatnonzero(flat != 0)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, iq=iq, bin_index=nz.astype(np.int32), bin_value=flat[nz],
                            bins_shape=np.array(bins.shape, np.int32),
                            digest=np.array(input_digest(wl)))
        print(f"{name}: iq {iq.shape}, {len(nz)} non-zero bins -> {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
