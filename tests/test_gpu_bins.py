"""Fixed-point LDS bins (parameter bin_accumulation: auto, fixed or float)
against float bins and the FP64 oracle.

Fixed-point bins add each contribution w*a, rounded to a grid of 2^-k
(amp_scale = 2^k, chosen per launch from the dataset's largest amplitude and
size), as 64-bit integers; the sums are exact, so the frame differs from the
float-bin frame only by float rounding (~1e-7) and stays at the oracle's bar.
What pins the choice of k and the kernel's conversion (|v| < 2^51 rounded with
a double addition): amplitudes scaled by 1e30 and 1e-30 must scale the frame
(a k that overflows or underflows the grid fails), a phantom with a strong
scatterer 1e6 above the rest must keep every line at the oracle's bar (a grid
too coarse for the weak lines fails), and a range the grid cannot resolve
(1e13) must fall back to float bins, as must a non-finite amplitude."""
import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "gauss_sector": lambda: W.fixed_sector_2d(20000, 16, seed=61),
    "gauss_sector_phase": lambda: W.fixed_sector_2d(20000, 16, seed=62, phase_delay=True),
    "lut_volume": lambda: W.volume_lut(30000, n_az=6, n_el=6, seed=63),
    "lut_volume_phase": lambda: W.volume_lut(30000, n_az=4, n_el=4, seed=64, phase_delay=True),
    "spline": lambda: W.spline_mmode(8000, 12, seed=65),
    "spline_equal_times": lambda: W.spline_mmode(8000, 12, seed=66, all_timestamps_equal=True),
}


def gpu_frame(wl, bins="fixed", **params):
    from openbcsim_amd.pyrfsim import RfSimulator
    g = RfSimulator("gpu")
    wl.apply(g)
    g.set_parameter("bin_accumulation", bins)
    for k, v in params.items():
        g.set_parameter(k, str(v))
    iq = g.simulate_frame()
    plan = g.get_parameter("launch_plan")
    g.close()
    return iq, plan


def oracle_frame(wl):
    """The FP64 oracle under the GPU's default weight-cutoff rule (1e-8 of the
    profile's peak), so that a strong scatterer's culled tails do not count as
    differences on the weak lines."""
    o = Oracle()
    wl.apply(o)
    o.set_parameter("weight_cutoff", "1e-8")
    return o.simulate_frame()


def per_line_errors(a, ref):
    return np.array([rel_l2(a[i], ref[i]) for i in range(ref.shape[0])])


@pytest.mark.parametrize("name", sorted(CASES))
def test_fixed_bins_match_float_bins_and_oracle(name):
    wl = CASES[name]()
    print(f"[case] {name}", flush=True)
    fx, plan_fx = gpu_frame(wl, "fixed")
    fl, plan_fl = gpu_frame(wl, "float")
    ref = oracle_frame(wl)
    print(f"[result] {name}: fixed vs float {rel_l2(fx, fl):.2e}, fixed vs oracle {rel_l2(fx, ref):.2e}, "
          f"float vs oracle {rel_l2(fl, ref):.2e}", flush=True)
    assert "bins=fixed" in plan_fx and "bins=float" not in plan_fx
    assert "bins=float" in plan_fl and "bins=fixed" not in plan_fl
    assert rel_l2(fx, fl) <= 2e-6
    assert rel_l2(fx, ref) <= 1e-5
    assert per_line_errors(fx, ref).max() <= 1e-5


def test_fixed_bins_are_order_independent():
    """With one chunk slice per block the bins leave LDS with plain stores, and
    integer sums do not depend on the order of the additions: two frames of the
    same simulator are equal bit for bit."""
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.volume_lut(30000, n_az=6, n_el=6, seed=67)
    g = RfSimulator("gpu")
    wl.apply(g)
    g.set_parameter("tile_slices", "1")
    g.set_parameter("bin_accumulation", "fixed")
    g.simulate_frame()
    a = g.get_projection().copy()
    g.simulate_frame()
    b = g.get_projection().copy()
    plan = g.get_parameter("launch_plan")
    g.close()
    print(f"[result] plan {plan.strip()!r}; differing bins {int(np.count_nonzero(a != b))}", flush=True)
    assert "slices=1" in plan and "bins=fixed" in plan
    assert np.array_equal(a, b)


@pytest.mark.parametrize("scale", [1e30, 1e-30])
def test_amplitude_scale_extremes(scale):
    wl = W.fixed_sector_2d(20000, 16, seed=68)
    base, _ = gpu_frame(wl, "fixed")
    wl.fixed = [wl.fixed[0].copy()]
    wl.fixed[0][:, 3] = (wl.fixed[0][:, 3].astype(np.float64) * scale).astype(np.float32)
    fx, plan = gpu_frame(wl, "fixed")
    ref = oracle_frame(wl)
    e = rel_l2(fx, ref)
    print(f"[result] amplitudes x{scale:g}: vs oracle {e:.2e}, vs scaled base {rel_l2(fx, base * scale):.2e}",
          flush=True)
    assert "bins=fixed" in plan
    assert e <= 1e-5
    assert rel_l2(fx, base * scale) <= 1e-5


def strong_scatterer(wl, ratio):
    """The phantom with one scatterer on line 0's axis raised to `ratio` times the
    largest amplitude of the rest."""
    pts = wl.fixed[0].copy()
    o, d = wl.scan[0][0], wl.scan[1][0]
    i = int(np.argmax(pts[:, 3]))
    pts[i, :3] = (o + 0.04 * d).astype(np.float32)
    pts[i, 3] = np.float32(ratio * float(np.abs(pts[:, 3]).max()))
    wl.fixed = [pts]
    return wl


def test_high_dynamic_range_keeps_weak_lines():
    """A scatterer 1e6 above the rest: every line, those that see only the weak
    scatterers included, stays at the oracle's bar."""
    wl = strong_scatterer(W.fixed_sector_2d(20000, 16, seed=69), 1e6)
    fx, plan = gpu_frame(wl, "fixed")
    ref = oracle_frame(wl)
    errs = per_line_errors(fx, ref)
    print(f"[result] ratio 1e6: {plan.strip()!r}; worst line vs oracle {errs.max():.2e}", flush=True)
    assert "bins=fixed" in plan
    assert errs.max() <= 1e-5


def test_unresolvable_range_falls_back_to_float():
    wl = strong_scatterer(W.fixed_sector_2d(20000, 16, seed=70), 1e13)
    fx, plan = gpu_frame(wl, "fixed")
    ref = oracle_frame(wl)
    errs = per_line_errors(fx, ref)
    print(f"[result] ratio 1e13: {plan.strip()!r}; worst line vs oracle {errs.max():.2e}", flush=True)
    assert "bins=float" in plan
    assert errs.max() <= 1e-4


def test_non_finite_amplitude_uses_float_bins():
    wl = W.fixed_sector_2d(4000, 8, seed=71)
    wl.fixed = [wl.fixed[0].copy()]
    wl.fixed[0][0, 3] = np.inf
    fx, plan = gpu_frame(wl, "fixed")
    print(f"[result] non-finite amplitude: {plan.strip()!r}; finite samples {np.isfinite(fx).mean():.3f}",
          flush=True)
    assert "bins=float" in plan
    assert not np.all(np.isfinite(fx))  # the infinite amplitude reaches the IQ, as in float arithmetic


def test_bin_accumulation_parameter():
    from openbcsim_amd.pyrfsim import RfSimulator
    g = RfSimulator("gpu")
    assert g.get_parameter("bin_accumulation") == "auto"
    g.set_parameter("bin_accumulation", "float")
    assert g.get_parameter("bin_accumulation") == "float"
    with pytest.raises(Exception):
        g.set_parameter("bin_accumulation", "double")
    g.close()


def test_automatic_choice():
    """auto: fixed point for the Gaussian profile, float for a LUT; with tuning on,
    the tuner times both and records the trials."""
    _, plan_g = gpu_frame(W.fixed_sector_2d(4000, 8, seed=72), "auto")
    _, plan_l = gpu_frame(W.volume_lut(6000, n_az=3, n_el=3, seed=73), "auto")
    print(f"[result] gaussian {plan_g.strip()!r}; lut {plan_l.strip()!r}", flush=True)
    assert "bins=fixed" in plan_g and "bins=float" in plan_l
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.fixed_sector_2d(30000, 24, seed=74)
    g = RfSimulator("gpu")
    wl.apply(g)
    g.set_parameter("autotune", "on")
    iq = g.simulate_frame()
    report = g.get_parameter("autotune_report")
    g.close()
    line = [r for r in report.splitlines() if "/gauss/" in r and "L2^5" in r]
    print(f"[result] tuner: {line}", flush=True)
    assert line and ",fixed]=" in line[0] and ",float]=" in line[0]
    assert rel_l2(iq, oracle_frame(wl)) <= 1e-5


def test_full_phase_volume_tuned_shape_matches_default():
    """The shape the tuner picks for the full 1M-point volume with phase delay
    (fixed-point complex bins, 4 radial windows of 4 lines, one-word chunk-list
    entries) against the default shape (float bins, whole 2-line tiles): equal
    up to float summation order."""
    wl = W.volume_lut(phase_delay=True)
    ref, plan_ref = gpu_frame(wl, "auto")
    iq, plan = gpu_frame(wl, "fixed", tile_windows=4, tile_lines=4)
    e = rel_l2(iq, ref)
    print(f"[result] tuned shape {plan.strip()!r} vs default {plan_ref.strip()!r}: rel_l2 {e:.2e}", flush=True)
    assert "windows=4" in plan and "lines=4" in plan and "bins=fixed" in plan and "packed" in plan
    assert "bins=float" in plan_ref
    assert e < 1e-6
