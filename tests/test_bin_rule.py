"""The float bin rule of the projection kernels without phase delay equals the
double rule everywhere.

radial_bin (proj_math.h) evaluates the reference's floor(r*scale/denom + 0.5) in
double (CpuAlgorithm.cpp:75 for fixed points, :161-162 for splines), with an
exact division where the quotient lies within 1e-9 of a bin edge.
radial_bin_nophase forms r*scale/denom + 0.5 as an unevaluated sum of floats
(exact product error, two-sum) and takes the bin from it unless its fraction
lies within 2.5e-7 of 0 or 1, where it falls back to the double rule.  Both are restated here operation for operation (numpy float32
arithmetic rounds like the device's, which compiles this code without
contraction) and compared on random radii and on the float radii nearest every
bin edge of a line and the 16 on either side of each -- the inputs where a too
narrow tolerance would show -- and on radii far outside the line.  The GPU test tiled-vs-direct compares the two
kernels' bins on whole frames."""
import numpy as np
import pytest

f32 = np.float32


@np.errstate(invalid="ignore", over="ignore")
def double_rule(r, scale, denom, num_rf):
    q = r.astype(np.float64) * scale
    ti = q * (1.0 / denom)
    t = ti + 0.5
    fl = np.floor(t)
    fr = t - fl
    edge = (fr < 1e-9) | (fr > 1.0 - 1e-9)
    t = np.where(edge, q / denom + 0.5, t)
    fl = np.floor(t)
    ok = (t >= 0.0) & (t < num_rf)
    return np.where(ok, fl, -1).astype(np.int64)


@np.errstate(invalid="ignore", over="ignore")
def float_rule(r, scale, denom, num_rf):
    """radial_bin_nophase: t = r*K + 0.5 as an unevaluated float sum (the product's
    exact error as the device's fused multiply-add gives it -- exactly
    representable, formed here in double), then the fraction test."""
    k = scale / denom
    k_hi = f32(k)
    k_lo = f32(k - np.float64(k_hi))
    r = r.astype(f32)
    p = (r * k_hi).astype(f32)
    e = (r.astype(np.float64) * np.float64(k_hi) - p.astype(np.float64)).astype(f32)
    s = (e + (r * k_lo).astype(f32)).astype(f32)
    t_hi = (p + f32(0.5)).astype(f32)
    bb = (t_hi - p).astype(f32)
    err = ((p - (t_hi - bb)) + (f32(0.5) - bb)).astype(f32)
    fl = np.floor(t_hi)
    f = ((t_hi - fl) + (err + s)).astype(f32)
    kk = np.floor(f)
    fr = (f - kk).astype(f32)
    clear = (fr > f32(2.5e-7)) & (fr < f32(1.0) - f32(2.5e-7))
    b = fl + kk
    fast = np.where((b >= 0) & (b < f32(num_rf)), b, -1).astype(np.int64)
    return np.where(clear, fast, double_rule(r, scale, denom, num_rf)), clear


def edge_radii(scale, denom, num_rf, around=16):
    n = np.arange(-1, num_rf + 2, dtype=np.float64)
    edges = ((n - 0.5) * denom / scale).astype(f32)
    out = [edges]
    up, down = edges.copy(), edges.copy()
    for _ in range(around):
        up = np.nextafter(up, f32(np.inf))
        down = np.nextafter(down, f32(-np.inf))
        out += [up, down]
    return np.concatenate(out)


@pytest.mark.parametrize("fs,c", [(50e6, 1540.0), (40e6, 1500.0), (33.3e6, 1540.0), (100e6, 1480.0)])
@pytest.mark.parametrize("rule", ["fixed", "spline"])
def test_float_rule_equals_double_rule(fs, c, rule):
    if rule == "fixed":  # r*2.0*fs/c
        scale, denom = 2.0 * fs, c
    else:  # r*2.0/(c*float(1/fs))
        scale, denom = 2.0, float(f32(c) * f32(1.0 / fs))
    line = 0.15
    num_rf = int(np.floor(2.0 * line * fs / c + 0.5))
    rng = np.random.default_rng(int(fs) % 1000)
    r = np.concatenate([rng.uniform(-0.01, 1.3 * line, 2_000_000).astype(f32), edge_radii(scale, denom, num_rf),
                        -edge_radii(scale, denom, 4), rng.uniform(-1e7, 1e7, 10000).astype(f32),
                        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 3e38, -3e38], f32)])
    want = double_rule(r, scale, denom, num_rf)
    got, clear = float_rule(r, scale, denom, num_rf)
    assert np.array_equal(got, want)
    # the float path decides all but about one random radius in a million
    assert clear[:2_000_000].mean() > 1.0 - 5e-6


def test_narrow_tolerance_would_fail():
    """The check can fail: a float rule without the fallback misplaces radii next to
    bin edges."""
    fs, c = 50e6, 1540.0
    scale, denom = 2.0 * fs, c
    num_rf = 9740
    r = edge_radii(scale, denom, num_rf)
    factor = f32(scale * (1.0 / denom))
    t = (r * factor).astype(f32) + f32(0.5)
    naive = np.where((t >= 0) & (t < num_rf), np.floor(t), -1).astype(np.int64)
    assert not np.array_equal(naive, double_rule(r, scale, denom, num_rf))
