"""Pins for the FP64 oracle (CPU only).

Every oracle function is checked against something other than its own formula:
the known answers in the reference's unit tests, independent library routines
(numpy FFT, scipy B-splines), brute force, closed forms, invariants, and the
golden frames written by the reference's own compiled CPU algorithm.
"""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline

from backends import Oracle, oracle_lib, rel_l2
from openbcsim_amd import workloads as W

_dp = __import__("ctypes").POINTER(__import__("ctypes").c_double)
_fp = __import__("ctypes").POINTER(__import__("ctypes").c_float)


def _d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(_dp)


# --- Hilbert mask: src/core/unittest/test_hilbert_mask.cpp:14-34 -------------------------
@pytest.mark.parametrize("n,expected", [
    (7, [1, 2, 2, 2, 0, 0, 0]),
    (8, [1, 2, 2, 2, 1, 0, 0, 0]),
    (11, [1, 2, 2, 2, 2, 2, 0, 0, 0, 0, 0]),
    (18, [1, 2, 2, 2, 2, 2, 2, 2, 2, 1, 0, 0, 0, 0, 0, 0, 0, 0]),
])
def test_hilbert_mask_known_answers(n, expected):
    out = np.zeros(n)
    oracle_lib().oracle_hilbert_mask(n, out.ctypes.data_as(_dp))
    assert out.tolist() == expected


# --- FFT length: src/core/unittest/test_fft.cpp:77-90 ------------------------------------
def test_next_pow2_known_answers():
    lib = oracle_lib()
    table = {1: 1, 2: 2, 3: 4, 4: 4, 5: 8, 7: 8, 8: 8, 15: 16, 16: 16, 32767: 32768,
             65535: 65536, 65536: 65536}
    for n, p in table.items():
        assert lib.oracle_next_pow2(n) == p


def test_next_pow2_matches_reference_formula():
    # fft.cpp:80-82 computes pow(2, ceil(log(n)/log(2))); identical for every n the path uses
    lib = oracle_lib()
    for n in list(range(1, 5000)) + [2 ** k + d for k in range(12, 21) for d in (-1, 0, 1)]:
        ref = int(math.pow(2, math.ceil(math.log(n) / math.log(2))))
        assert lib.oracle_next_pow2(n) == ref, n


# --- FFT convolution: src/core/unittest/test_fft.cpp:55-75 -------------------------------
def _fft_conv(x, y):
    xa, xp = _d(x)
    ya, yp = _d(y)
    out = np.zeros(len(x) + len(y) - 1)
    oracle_lib().oracle_fft_conv(xp, len(x), yp, len(y), out.ctypes.data_as(_dp))
    return out


def test_fft_conv_known_answer_1():
    assert np.allclose(_fft_conv([1.0, 2.0, 3.0], [1.0]), [1.0, 2.0, 3.0], rtol=1e-4)


def test_fft_conv_known_answer_2():
    x = [1.0, 2.0, 3.0, -0.23, 0.001, 32.3, 4.0]
    h = [-0.33, 0.9, -0.002, 1.1, 2.3]
    desired = [-0.33, 0.24, 0.808, 3.8719, 4.28667, -2.75764, 34.396998, 3.0075, 35.5243, 78.69, 9.2]
    # the reference checks 0.01 % (BOOST_CHECK_CLOSE(..., 0.01f))
    assert np.allclose(_fft_conv(x, h), desired, rtol=1e-4)


def test_direct_conv_known_answer():
    # src/utils/unittest/test_SignalProcessing.cpp:16-27
    got = _fft_conv([1, -1, 0, 10, 3], [3, 9, -2])
    assert np.allclose(got, [3, 6, -11, 32, 99, 7, -6], atol=1e-9)


@pytest.mark.parametrize("n", [1, 2, 4, 16, 64, 1024, 8192, 16384])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    buf = np.empty(2 * n)
    buf[0::2], buf[1::2] = x.real, x.imag
    oracle_lib().oracle_fft(buf.ctypes.data_as(_dp), n, 0)
    got = buf[0::2] + 1j * buf[1::2]
    assert rel_l2(got, np.fft.fft(x)) < 1e-13
    buf[0::2], buf[1::2] = x.real, x.imag
    oracle_lib().oracle_fft(buf.ctypes.data_as(_dp), n, 1)
    got = buf[0::2] + 1j * buf[1::2]
    assert rel_l2(got, np.fft.ifft(x) * n) < 1e-13


def test_fft_matches_brute_force_dft():
    n = 32
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    buf = np.empty(2 * n)
    buf[0::2], buf[1::2] = x.real, x.imag
    oracle_lib().oracle_fft(buf.ctypes.data_as(_dp), n, 0)
    assert rel_l2(buf[0::2] + 1j * buf[1::2], dft) < 1e-13


# --- sample counts ------------------------------------------------------------------------
def test_num_rf_samples_known_values():
    lib = oracle_lib()
    assert lib.oracle_num_rf_samples(1540.0, 0.12, 50e6) == 7792
    assert lib.oracle_num_rf_samples(1540.0, 0.15, 50e6) == 9740
    assert lib.oracle_num_rf_samples(1540.0, 0.04, 50e6) == 2597


def test_gaussian_excitation_shape():
    exc, center = W.make_gaussian_excitation()
    assert len(exc) == 55 and center == 27
    assert exc[center] == pytest.approx(1.0)
    assert np.allclose(exc, exc[::-1], atol=1e-6)


# --- B-splines: bspline.hpp:45-91 against scipy's independent evaluation ---------------------
def _basis(j, p, t, knots):
    k, kp = _d(knots)
    return oracle_lib().oracle_bspline_basis(j, p, t, kp, len(k))


@pytest.mark.parametrize("p,m", [(1, 4), (2, 7), (3, 10), (4, 12)])
def test_bspline_basis_matches_scipy(p, m):
    knots = W.uniform_knot_vector(m, p).astype(np.float64)
    for t in np.linspace(0.0, 0.999, 41):
        vals = np.array([_basis(j, p, t, knots) for j in range(m)])
        ref = np.array([BSpline(knots, np.eye(m)[j], p)(t) for j in range(m)])
        assert np.allclose(vals, ref, atol=1e-12), t
        assert abs(vals.sum() - 1.0) < 1e-12  # partition of unity


def test_bspline_float_basis_close_to_double():
    knots = W.uniform_knot_vector(10, 3)
    kf = np.ascontiguousarray(knots, np.float32)
    lib = oracle_lib()
    for t in np.linspace(0.0, 0.999, 17).astype(np.float32):
        for j in range(10):
            f = lib.oracle_bspline_basis_f32(j, 3, float(t), kf.ctypes.data_as(_fp), len(kf))
            d = _basis(j, 3, float(t), knots.astype(np.float64))
            assert abs(f - d) < 1e-6


def test_knot_interval():
    knots = np.ascontiguousarray(W.uniform_knot_vector(10, 3), np.float32)
    lib = oracle_lib()
    kp = knots.ctypes.data_as(_fp)
    assert lib.oracle_knot_interval(kp, len(knots), 0.0) == 3
    assert lib.oracle_knot_interval(kp, len(knots), 0.5) == 6
    assert lib.oracle_knot_interval(kp, len(knots), 1.0) == -1  # bspline.hpp:81 throws


# --- beam profiles ---------------------------------------------------------------------------
def test_lut_affine_table_interpolates_exactly():
    o = Oracle()
    nr, nl, ne = 5, 4, 6
    r = np.linspace(0.0, 0.1, nr)
    l = np.linspace(-0.01, 0.01, nl)
    e = np.linspace(-0.02, 0.02, ne)
    f = lambda rr, ll, ee: 0.3 + 2.0 * rr - 5.0 * ll + 7.0 * ee
    tab = f(r[:, None, None], l[None, :, None], e[None, None, :]).astype(np.float32)
    o.set_lut_beam_profile(0.0, 0.1, -0.01, 0.01, -0.02, 0.02, tab)
    rng = np.random.default_rng(0)
    for _ in range(200):
        rr, ll, ee = rng.uniform(0.0, 0.0999), rng.uniform(-0.0099, 0.0099), rng.uniform(-0.0199, 0.0199)
        assert o.lut_sample(rr, ll, ee) == pytest.approx(f(rr, ll, ee), abs=2e-6)
    # outside the table, and on the upper edge where the upper neighbour is missing: zero
    assert o.lut_sample(0.11, 0.0, 0.0) == 0.0
    assert o.lut_sample(0.05, 0.02, 0.0) == 0.0
    assert o.lut_sample(float(np.float32(0.1)), 0.0, 0.0) == 0.0  # r1 = n_rad is out of range
    assert o.lut_sample(0.1, 0.0, 0.0) > 0.0  # double 0.1 is just below the float grid end
    # BeamProfile.cpp:87-88 truncates toward zero: just below the lower edge reads the edge
    assert o.lut_sample(-0.001, 0.0, 0.0) == pytest.approx(f(0.0, 0.0, 0.0), abs=1e-3)


def _single_scatterer_frame(pos, amp=1.0, phase_delay=False, decimation=1, arc=True, sigma=1e-3):
    wl = W.fixed_sector_2d(1, 1, line_length=0.04, phase_delay=phase_delay, arc=arc,
                           decimation=decimation, sigma=sigma)
    o, d, l, t, length = W.linear_scan(1, 0.0, 0.04) if False else (
        np.zeros((1, 3), np.float32), np.array([[0, 0, 1]], np.float32),
        np.array([[1, 0, 0]], np.float32), np.zeros(1, np.float32), 0.04)
    wl.scan = (o, d, l, t, length)
    wl.fixed = [np.array([[pos[0], pos[1], pos[2], amp]], np.float32)]
    orc = Oracle()
    wl.apply(orc)
    return wl, orc.simulate_frame()[0]


def test_single_on_axis_scatterer_is_the_analytic_pulse():
    """Closed form: one scatterer on the beam axis at depth z gives
    IQ[j] = a * h[center + jD - b] * exp(-2 pi i nfd jD), with h the analytic
    excitation (numpy FFT, Hilbert mask) and b = round(2 fs z / c)."""
    z, amp, D = 0.02, float(np.float32(0.7)), 2
    wl, iq = _single_scatterer_frame((0.0, 0.0, z), amp, decimation=D)
    N = wl.fft_len()
    exc = wl.excitation.astype(np.float64)
    Hx = np.fft.fft(exc, N)
    mask = np.zeros(N)
    mask[0] = 1
    mask[1:N // 2] = 2
    mask[N // 2] = 1
    h = np.fft.ifft(Hx * mask)
    b = int(math.floor(float(np.float32(z)) * 2.0 * float(np.float32(W.FS)) / float(np.float32(1540.0)) + 0.5))
    nfd = float(np.float32(np.float32(W.FC) / np.float32(W.FS)))
    j = np.arange(iq.shape[0])
    n = j * D
    idx = wl.center_index + n - b
    # the analytic excitation is N-periodic (its Hilbert part has unbounded support)
    expect = amp * h[idx % N] * np.exp(-2j * np.pi * nfd * n)
    assert rel_l2(iq, expect) < 1e-12


def test_gaussian_weight_closed_form():
    # a scatterer displaced laterally by x from the axis is weighted exp(-x^2/(2 sigma^2))
    _, on = _single_scatterer_frame((0.0, 0.0, 0.02))
    x = 0.8e-3
    _, off = _single_scatterer_frame((x, 0.0, 0.02), arc=False)
    w = math.exp(-(float(np.float32(x)) ** 2) / (2 * float(np.float32(1e-3)) ** 2))
    assert np.allclose(off, w * on, rtol=1e-9, atol=1e-12)


def test_linearity_superposition_decimation():
    wl = W.fixed_sector_2d(3000, 5, seed=3, line_length=0.04, phase_delay=True)
    o = Oracle()
    wl.apply(o)
    base = o.simulate_frame()
    # amplitude scaling by 2 is exact in binary
    wl2 = W.fixed_sector_2d(3000, 5, seed=3, line_length=0.04, phase_delay=True)
    wl2.fixed[0][:, 3] *= 2.0
    o2 = Oracle()
    wl2.apply(o2)
    assert rel_l2(o2.simulate_frame(), 2 * base) < 1e-13
    # splitting the dataset in two gives the same frame
    wl3 = W.fixed_sector_2d(3000, 5, seed=3, line_length=0.04, phase_delay=True)
    pts = wl3.fixed[0]
    wl3.fixed = [pts[:1234], pts[1234:]]
    o3 = Oracle()
    wl3.apply(o3)
    assert rel_l2(o3.simulate_frame(), base) < 1e-13
    # decimation by 3 picks every third sample of the full-rate frame
    wl4 = W.fixed_sector_2d(3000, 5, seed=3, line_length=0.04, phase_delay=True, decimation=3)
    o4 = Oracle()
    wl4.apply(o4)
    assert np.array_equal(o4.simulate_frame(), base[:, ::3])


def test_constant_spline_equals_fixed_scatterer():
    """A spline whose control points all coincide stays at that point for every t."""
    fx = W.fixed_sector_2d(500, 4, seed=8, line_length=0.04)
    sp = W.spline_mmode(500, 4, seed=9, line_length=0.04, decimation=1)
    pts = fx.fixed[0]
    cps = np.repeat(pts[:, None, :3], 10, axis=1).astype(np.float32)
    sp.splines[0].control_points = cps
    sp.splines[0].amplitudes = pts[:, 3].copy()
    sp.scan = fx.scan[:3] + (sp.scan[3], fx.scan[4])
    a, b = Oracle(), Oracle()
    fx.apply(a)
    sp.apply(b)
    # The float sum of cp*B_i (CpuAlgorithm.cpp:137-140) reproduces cp only to an ulp or so,
    # which moves the Gaussian weight by ~1e-7.  The spline bin rule divides by c*float(1/fs)
    # (:161-162), the fixed rule by c/fs; a point on a half-sample boundary could differ, and
    # none does at this seed: the bins carry the same support.
    pa, pb = a.project(), b.project()
    assert np.array_equal(np.abs(pa) > 1e-20, np.abs(pb) > 1e-20)
    assert rel_l2(b.simulate_frame(), a.simulate_frame()) < 1e-6


def test_arc_projection_irrelevant_on_axis():
    _, a = _single_scatterer_frame((0.0, 0.0, 0.03), arc=True)
    _, b = _single_scatterer_frame((0.0, 0.0, 0.03), arc=False)
    assert np.array_equal(a, b)


def test_invalid_configuration_errors():
    o = Oracle()
    with pytest.raises(RuntimeError, match="Excitation must be configured"):
        o.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.04, np.array([[1.0, 0, 0]]), np.zeros(1))
    exc, c = W.make_gaussian_excitation()
    o.set_excitation(exc, c, W.FS, W.FC)
    with pytest.raises(RuntimeError, match="orthonormal"):
        o.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.04, np.array([[1.0, 0.1, 0]]), np.zeros(1))
    with pytest.raises(RuntimeError, match="illegal radial decimation"):
        o.set_parameter("radial_decimation", "0")
    with pytest.raises(RuntimeError, match="illegal parameter name"):
        o.set_parameter("no_such_key", "1")
