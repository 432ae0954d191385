"""Seeded input generators: excitation pulse, scan sequences and phantoms.

These restate the reference's input builders so every backend (the HIP
library, the FP64 oracle and the compiled reference CPU algorithm) is fed the
same float32 arrays:

* ``make_gaussian_excitation`` follows ``MakeGaussianExcitation`` and
  ``GaussianPulse`` (src/utils/GaussPulse.cpp:33-70), including its float
  evaluation, so the pulse for fc=2.5 MHz, bw=0.5, fs=50 MHz has 55 samples
  with centre index 27.
* ``sector_scan`` and ``linear_scan`` follow ``CreateScanSequence``
  (src/utils/BCSimConvenience.cpp:117-175): rotation about y built in double,
  then cast to float.
* ``pyramid_scan`` is the 3D volume scan of the 1M-scatterer workload.
* ``uniform_knot_vector`` follows ``uniform_regular_knot_vector``
  (src/core/bspline.hpp:93-111).

The workload builders (``fixed_sector_2d``, ``spline_mmode``, ``volume_lut``,
``batched_jobs``) produce the five benchmark configurations at full size or at
reduced sizes for parity tests.  Random numbers come from numpy's PCG64 with an
explicit seed.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

SOUND_SPEED = 1540.0
FS = 50e6
FC = 2.5e6
BW = 0.5


def gaussian_pulse_time_limit(fc: float, bw: float, bwr: float = -6.0, tpr: float = -60.0) -> float:
    """GaussianPulseTimeLimits (GaussPulse.cpp:33-38), double precision."""
    e = math.exp(1.0)
    f_cutoff = bw * fc
    return math.sqrt(-tpr / (20.0 * (-math.pi ** 2 * f_cutoff * f_cutoff
                                     / math.log(10.0 ** (bwr / 20.0))) * math.log10(e)))


def make_gaussian_excitation(fc: float = FC, bw: float = BW, fs: float = FS) -> Tuple[np.ndarray, int]:
    """MakeGaussianExcitation<float> (GaussPulse.cpp:51-70).

    Returns (samples float32, center_index)."""
    f32 = np.float32
    limit = gaussian_pulse_time_limit(float(f32(fc)), float(f32(bw)))
    num = int(2.0 * limit * float(f32(fs)))
    pi_f = f32(4.0 * math.atan(1.0))
    fc_f, bw_f, fs_f = f32(fc), f32(bw), f32(fs)
    f_cut = f32(bw_f * fc_f)
    denom = math.log(10.0 ** (-6.0 / 20.0))
    out = np.empty(num, dtype=np.float32)
    for i in range(num):
        t = f32(f32(i - num // 2) / fs_f)
        ftp = f32(f_cut * t)
        a = f32(f32(f32(pi_f * pi_f) * ftp) * ftp)
        env = math.exp(float(a) / denom)
        ph = f32(f32(f32(f32(2.0) * pi_f) * fc_f) * t)
        out[i] = f32(env * float(np.cos(ph)))
    return out, num // 2


def sector_scan(num_lines: int, width: float, depth: float, tilt: float = 0.0,
                timestamps: Optional[np.ndarray] = None):
    """Sector scan about the y axis, origin at 0 (BCSimConvenience.cpp:117-146)."""
    f32 = np.float32
    origins = np.zeros((num_lines, 3), np.float32)
    dirs = np.zeros((num_lines, 3), np.float32)
    lats = np.zeros((num_lines, 3), np.float32)
    for k in range(num_lines):
        step = f32(f32(k) * f32(width)) / f32(num_lines - 1) if num_lines > 1 else f32(0)
        angle = f32(f32(f32(-0.5) * f32(width)) + f32(tilt)) + f32(step)
        a = float(angle)
        dirs[k] = (math.sin(a), 0.0, math.cos(a))
        lats[k] = (math.cos(a), 0.0, -math.sin(a))
    ts = np.zeros(num_lines, np.float32) if timestamps is None else np.asarray(timestamps, np.float32)
    return origins, dirs, lats, ts, float(depth)


def linear_scan(num_lines: int, width: float, range_max: float,
                timestamps: Optional[np.ndarray] = None):
    """Linear scan along x, beams along z (BCSimConvenience.cpp:148-175)."""
    f32 = np.float32
    origins = np.zeros((num_lines, 3), np.float32)
    for k in range(num_lines):
        origins[k, 0] = f32(-0.5) * f32(width) + f32(k) * f32(width) / f32(num_lines - 1)
    dirs = np.tile(np.array([0, 0, 1], np.float32), (num_lines, 1))
    lats = np.tile(np.array([1, 0, 0], np.float32), (num_lines, 1))
    ts = np.zeros(num_lines, np.float32) if timestamps is None else np.asarray(timestamps, np.float32)
    return origins, dirs, lats, ts, float(range_max)


def pyramid_scan(n_az: int, n_el: int, half_angle_az: float, half_angle_el: float, depth: float,
                 timestamps: Optional[np.ndarray] = None):
    """Pyramidal volume scan from the origin: direction R_y(az) R_x(-el) z,
    lateral R_y(az) x.  Built in double, cast to float."""
    L = n_az * n_el
    origins = np.zeros((L, 3), np.float32)
    dirs = np.zeros((L, 3), np.float32)
    lats = np.zeros((L, 3), np.float32)
    azs = np.linspace(-half_angle_az, half_angle_az, n_az)
    els = np.linspace(-half_angle_el, half_angle_el, n_el)
    k = 0
    for el in els:
        for az in azs:
            dirs[k] = (math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az))
            lats[k] = (math.cos(az), 0.0, -math.sin(az))
            k += 1
    ts = np.zeros(L, np.float32) if timestamps is None else np.asarray(timestamps, np.float32)
    return origins, dirs, lats, ts, float(depth)


def uniform_knot_vector(n: int, p: int, t0: float = 0.0, t1: float = 1.0) -> np.ndarray:
    """p+1-regular uniform knot vector for n control points (bspline.hpp:93-111)."""
    if n < p + 1:
        raise ValueError("Too small n for a uniform regular knot vector")
    num_middle = n + 1 - p
    mid = [t0 + i * (t1 - t0) / (num_middle - 1) for i in range(num_middle)]
    return np.array([t0] * p + mid + [t1] * p, np.float32)


def gaussian_lut(n_rad: int, n_lat: int, n_ele: int, r_range, l_range, e_range,
                 sigma0: float, sigma1: float) -> np.ndarray:
    """Beam-profile table [n_rad][n_lat][n_ele] (the order of
    utilities/make_lut.py:28-43): exp(-(l^2+e^2)/(2 sigma(r)^2)), sigma linear in r."""
    r = np.linspace(r_range[0], r_range[1], n_rad)
    l = np.linspace(l_range[0], l_range[1], n_lat)
    e = np.linspace(e_range[0], e_range[1], n_ele)
    sig = sigma0 + (sigma1 - sigma0) * (r - r_range[0]) / (r_range[1] - r_range[0])
    ll, ee = np.meshgrid(l, e, indexing="ij")
    tab = np.exp(-(ll[None] ** 2 + ee[None] ** 2) / (2.0 * sig[:, None, None] ** 2))
    return tab.astype(np.float32)


@dataclasses.dataclass
class SplineSet:
    degree: int
    knots: np.ndarray          # float32 [n_knots]
    control_points: np.ndarray  # float32 [N][M][3]
    amplitudes: np.ndarray     # float32 [N]


@dataclasses.dataclass
class Workload:
    """One simulation: everything the IAlgorithm setters take."""
    name: str
    excitation: np.ndarray
    center_index: int
    fs: float
    demod_freq: float
    scan: tuple                      # (origins, dirs, lats, timestamps, line_length)
    fixed: List[np.ndarray] = dataclasses.field(default_factory=list)   # each float32 [N][4]
    splines: List[SplineSet] = dataclasses.field(default_factory=list)
    profile: tuple = ("gaussian", 1e-3, 1e-3)   # or ("lut", (r0,r1,l0,l1,e0,e1), table)
    params: dict = dataclasses.field(default_factory=dict)
    all_timestamps_equal: bool = False

    @property
    def num_lines(self) -> int:
        return int(self.scan[0].shape[0])

    @property
    def num_scatterers(self) -> int:
        return sum(int(f.shape[0]) for f in self.fixed) + sum(int(s.amplitudes.shape[0]) for s in self.splines)

    def num_rf(self, sound_speed: float = SOUND_SPEED) -> int:
        f32 = np.float32
        c = f32(self.params.get("sound_speed", sound_speed))
        max_time = f32(2.0 * float(f32(self.scan[4])) / float(c))
        return int(math.floor(float(f32(f32(self.fs) * max_time)) + 0.5))

    def num_samples(self) -> int:
        d = int(self.params.get("radial_decimation", 1))
        return (self.num_rf() + d - 1) // d

    def fft_len(self) -> int:
        n = self.num_rf() + len(self.excitation) - 1
        return 1 << (n - 1).bit_length()

    def apply(self, sim) -> None:
        """Configure any simulator object exposing the IAlgorithm-style method set."""
        for k, v in self.params.items():
            sim.set_parameter(k, str(v))
        sim.set_excitation(self.excitation, self.center_index, self.fs, self.demod_freq)
        o, d, l, t, length = self.scan
        sim.set_scan_sequence(o, d, length, l, t, self.all_timestamps_equal)
        for f in self.fixed:
            sim.add_fixed_scatterers(f)
        for s in self.splines:
            sim.add_spline_scatterers(s.degree, s.knots, s.control_points, s.amplitudes)
        if self.profile[0] == "gaussian":
            sim.set_analytical_beam_profile(self.profile[1], self.profile[2])
        else:
            (r0, r1, l0, l1, e0, e1), table = self.profile[1], self.profile[2]
            sim.set_lut_beam_profile(r0, r1, l0, l1, e0, e1, table)


def _excitation():
    exc, center = make_gaussian_excitation(FC, BW, FS)
    return exc, center


def fixed_sector_2d(num_scatterers: int = 100_000, num_lines: int = 128, seed: int = 0,
                    line_length: float = 0.12, phase_delay: bool = False,
                    arc: bool = True, decimation: int = 1, sigma: float = 1e-3) -> Workload:
    """Plain fixed-scatterer 2D sector scan (100k points, 128 lines over +-35 deg)."""
    rng = np.random.default_rng(seed)
    pts = np.empty((num_scatterers, 4), np.float32)
    pts[:, 0] = rng.uniform(-0.03, 0.03, num_scatterers)
    pts[:, 1] = rng.uniform(-0.005, 0.005, num_scatterers)
    pts[:, 2] = rng.uniform(0.01, 0.12, num_scatterers)
    pts[:, 3] = rng.standard_normal(num_scatterers)
    exc, center = _excitation()
    scan = sector_scan(num_lines, math.radians(70.0), line_length)
    return Workload("fixed_sector_2d", exc, center, FS, FC, scan, fixed=[pts],
                    profile=("gaussian", sigma, sigma),
                    params={"sound_speed": SOUND_SPEED, "use_arc_projection": "on" if arc else "off",
                            "phase_delay": "on" if phase_delay else "off",
                            "radial_decimation": decimation})


def spline_mmode(num_scatterers: int = 200_000, num_lines: int = 256, seed: int = 1,
                 num_cp: int = 10, degree: int = 3, line_length: float = 0.12,
                 phase_delay: bool = False, decimation: int = 4, sigma: float = 1e-3,
                 all_timestamps_equal: bool = False) -> Workload:
    """Dynamic cubic-spline phantom, one time value per line (200k splines, 256 lines)."""
    rng = np.random.default_rng(seed)
    base = np.empty((num_scatterers, 1, 3), np.float32)
    base[:, 0, 0] = rng.uniform(-0.03, 0.03, num_scatterers)
    base[:, 0, 1] = rng.uniform(-0.005, 0.005, num_scatterers)
    base[:, 0, 2] = rng.uniform(0.01, 0.12, num_scatterers)
    cps = (base + rng.normal(0.0, 0.5e-3, (num_scatterers, num_cp, 3))).astype(np.float32)
    amps = rng.standard_normal(num_scatterers).astype(np.float32)
    knots = uniform_knot_vector(num_cp, degree, 0.0, 1.0)
    if all_timestamps_equal:
        ts = np.full(num_lines, rng.uniform(0.0, 1.0), np.float32)
    else:
        ts = rng.uniform(0.0, 1.0, num_lines).astype(np.float32)
    ts = np.minimum(ts, np.float32(np.nextafter(np.float32(1.0), np.float32(0.0))))
    exc, center = _excitation()
    scan = sector_scan(num_lines, math.radians(70.0), line_length, timestamps=ts)
    return Workload("spline_mmode", exc, center, FS, FC, scan,
                    splines=[SplineSet(degree, knots, cps, amps)],
                    profile=("gaussian", sigma, sigma),
                    params={"sound_speed": SOUND_SPEED, "use_arc_projection": "on",
                            "phase_delay": "on" if phase_delay else "off",
                            "radial_decimation": decimation},
                    all_timestamps_equal=all_timestamps_equal)


LUT_EXTENT = (0.0, 0.15, -0.01, 0.01, -0.01, 0.01)


def volume_lut(num_scatterers: int = 1_000_000, n_az: int = 64, n_el: int = 64, seed: int = 2,
               line_length: float = 0.15, lut_shape=(256, 64, 64), phase_delay: bool = False,
               decimation: int = 2) -> Workload:
    """3D volume scan with a trilinear LUT beam profile (1M points, 64x64 lines)."""
    rng = np.random.default_rng(seed)
    pts = np.empty((num_scatterers, 4), np.float32)
    pts[:, 0] = rng.uniform(-0.03, 0.03, num_scatterers)
    pts[:, 1] = rng.uniform(-0.03, 0.03, num_scatterers)
    pts[:, 2] = rng.uniform(0.01, 0.12, num_scatterers)
    pts[:, 3] = rng.standard_normal(num_scatterers)
    exc, center = _excitation()
    scan = pyramid_scan(n_az, n_el, math.radians(30.0), math.radians(30.0), line_length)
    r0, r1, l0, l1, e0, e1 = LUT_EXTENT
    table = gaussian_lut(*lut_shape, (r0, r1), (l0, l1), (e0, e1), 0.5e-3, 3e-3)
    return Workload("volume_lut", exc, center, FS, FC, scan, fixed=[pts],
                    profile=("lut", LUT_EXTENT, table),
                    params={"sound_speed": SOUND_SPEED, "use_arc_projection": "on",
                            "phase_delay": "on" if phase_delay else "off",
                            "radial_decimation": decimation})


FFT_LINE_LENGTHS = {2048: 0.028, 4096: 0.060, 8192: 0.12, 16384: 0.15}


def batched_jobs(num_jobs: int = 64, seed: int = 3, scale: float = 1.0) -> List[Workload]:
    """64 uneven jobs: scatterer counts log-uniform in [10k, 2M] (times ``scale``),
    fixed or cubic splines with 6-16 control points, 32-512 lines, FFT sizes
    2048-16384, analytic or LUT profile."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(num_jobs):
        n = int(round(math.exp(rng.uniform(math.log(10_000), math.log(2_000_000))) * scale))
        n = max(n, 16)
        lines = int(rng.integers(32, 513))
        nfft = int(rng.choice([2048, 4096, 8192, 16384]))
        length = FFT_LINE_LENGTHS[nfft]
        use_spline = bool(rng.integers(0, 2))
        use_lut = bool(rng.integers(0, 2))
        dec = int(rng.choice([1, 2, 4]))
        sub = int(rng.integers(0, 1 << 30))
        if use_spline:
            m = int(rng.integers(6, 17))
            wl = spline_mmode(n, lines, seed=sub, num_cp=m, line_length=length, decimation=dec)
        else:
            wl = fixed_sector_2d(n, lines, seed=sub, line_length=length, decimation=dec)
        if use_lut:
            r0, r1 = 0.0, length
            table = gaussian_lut(64, 32, 32, (r0, r1), (-0.01, 0.01), (-0.01, 0.01), 0.5e-3, 3e-3)
            wl.profile = ("lut", (r0, r1, -0.01, 0.01, -0.01, 0.01), table)
        wl.name = f"job{j}"
        jobs.append(wl)
    return jobs


def at_time(wl: Workload, geometry: tuple, t: float) -> Workload:
    """The frame a serving request (phantom ``wl``, scan ``geometry`` =
    (origins, dirs, lats, line_length), time ``t``) stands for: the phantom
    with that scan, every timestamp t."""
    o, d, l, length = geometry
    out = dataclasses.replace(wl)
    out.scan = (o, d, l, np.full(o.shape[0], np.float32(t), np.float32), float(length))
    out.all_timestamps_equal = True
    return out


def serving_mix(seed: int = 5, scale: float = 1.0):
    """16 phantoms and the scan geometries of the serving workload, drawn from the
    2D spline (M-mode-like), 3D LUT volume and uneven batched mixes:
      phantoms: 6 cubic-spline phantoms (200k, 10 control points), 4 fixed 2D
      sector phantoms (100k), 2 LUT volumes (1M), 4 uneven jobs of the batched
      mix -- scatterer counts times ``scale``;
      geometries: sector scans of 32-512 lines at the phantoms' depths and the
      64x64 pyramid (4096 lines).
    Returns (phantoms, geometries, pairs): ``pairs`` lists the (phantom,
    geometry) combinations requests draw from (a geometry whose depth fits the
    phantom's LUT extent)."""
    rng = np.random.default_rng(seed)
    phantoms: List[Workload] = []
    for k in range(6):
        phantoms.append(spline_mmode(max(16, int(200_000 * scale)), 64, seed=100 + k, decimation=4))
    for k in range(4):
        phantoms.append(fixed_sector_2d(max(16, int(100_000 * scale)), 64, seed=200 + k))
    for k in range(2):
        phantoms.append(volume_lut(max(16, int(1_000_000 * scale)), n_az=8, n_el=8, seed=300 + k))
    for wl in batched_jobs(4, seed=400, scale=scale):
        phantoms.append(wl)
    geometries = []
    for n in (32, 64, 128, 256, 512):
        o, d, l, _, _ = sector_scan(n, math.radians(70.0), 0.12)
        geometries.append((o, d, l, 0.12))
    o, d, l, _, _ = pyramid_scan(64, 64, math.radians(30.0), math.radians(30.0), 0.15)
    geometries.append((o, d, l, 0.15))
    pairs = []
    for p, wl in enumerate(phantoms):
        for g, (o, d, l, length) in enumerate(geometries):
            if wl.profile[0] == "lut" and length > wl.profile[1][1]:
                continue  # beyond the LUT's radial extent the profile is zero
            if o.shape[0] == 4096 and wl.name != "volume_lut":
                continue  # the pyramid scans the 3D phantoms
            pairs.append((p, g))
    rng.shuffle(pairs)
    return phantoms, geometries, pairs
