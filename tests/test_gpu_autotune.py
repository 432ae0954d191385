"""Launch-shape choices (FFT pass radix, tuned tile shapes) change float
summation order only: every choice matches the FP64 oracle at the parity bar
and the default choice to float precision."""
import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_sim():
    from openbcsim_amd.pyrfsim import RfSimulator
    return RfSimulator("gpu")


RADIX_CASES = {
    # FFT lengths 2048 (log2 11 = 3*3+2), 4096 (3*4), 8192 (3*4+1), 16384 (3*4+2); real and complex bins
    "fft2048_real": lambda: W.fixed_sector_2d(4000, 8, seed=71, line_length=0.028),
    "fft4096_complex": lambda: W.fixed_sector_2d(4000, 8, seed=72, line_length=0.06, phase_delay=True),
    "fft8192_real_dec3": lambda: W.fixed_sector_2d(4000, 8, seed=73, line_length=0.12, decimation=3),
    "fft16384_complex": lambda: W.volume_lut(20000, n_az=3, n_el=3, seed=74, phase_delay=True),
}


@pytest.mark.parametrize("name", sorted(RADIX_CASES))
def test_radix8_matches_radix4_and_oracle(name):
    wl = RADIX_CASES[name]()
    frames = {}
    for r in (4, 8):
        g = gpu_sim()
        g.set_parameter("fft_radix", str(r))
        wl.apply(g)
        frames[r] = g.simulate_frame()
        assert f"radix={r}" in g.get_parameter("launch_plan")
        g.close()
    o = Oracle()
    wl.apply(o)
    ref = o.simulate_frame()
    e8, e48 = rel_l2(frames[8], ref), rel_l2(frames[8], frames[4])
    print(f"[result] {name} N_fft={wl.fft_len()} radix8 vs oracle {e8:.2e} vs radix4 {e48:.2e}", flush=True)
    assert e8 <= 1e-4 and e48 < 1e-6


def test_autotuned_frames_match_default():
    """Tuning trials run on the frame's own bins before they are cleared: the
    tuned frame equals the untuned one, and the class is tuned once."""
    from openbcsim_amd.pyrfsim import RfSimulator
    cases = [W.fixed_sector_2d(30000, 48, seed=75, phase_delay=True),
             W.spline_mmode(8000, 32, seed=76),
             W.spline_mmode(8000, 32, seed=77, all_timestamps_equal=True),
             W.volume_lut(40000, n_az=5, n_el=5, seed=78)]
    for wl in cases:
        plain, tuned = gpu_sim(), gpu_sim()
        tuned.set_parameter("autotune", "on")
        wl.apply(plain)
        wl.apply(tuned)
        a = plain.simulate_frame()
        b = tuned.simulate_frame()
        report = tuned.get_parameter("autotune_report")
        assert "tile/" in report and "conv/" in report
        c = tuned.simulate_frame()  # cached shapes: no new trials
        assert tuned.get_parameter("autotune_report") == report
        err = max(rel_l2(b, a), rel_l2(c, a))
        print(f"[result] {wl.name} tuned vs default {err:.2e}; plan: "
              f"{tuned.get_parameter('launch_plan').strip()!r}", flush=True)
        assert err < 1e-6
        plain.close()
        tuned.close()


def test_autotune_keeps_noise_stream():
    """Trials do not advance the noise counter: frame k of a tuned simulator has
    the noise of frame k of an untuned one."""
    wl = W.fixed_sector_2d(5000, 8, seed=79)
    sims = []
    for tune in ("off", "on"):
        g = gpu_sim()
        g.set_parameter("autotune", tune)
        wl.apply(g)
        g.set_parameter("noise_amplitude", "0.5")
        g.set_parameter("noise_seed", "1234")
        sims.append(g)
    for _ in range(2):
        a, b = (g.simulate_frame() for g in sims)
        assert rel_l2(b, a) < 1e-6


def test_radix_parameter_errors():
    from openbcsim_amd.pyrfsim import BcsimError
    g = gpu_sim()
    with pytest.raises(BcsimError, match="fft_radix"):
        g.set_parameter("fft_radix", "5")
    g.set_parameter("fft_radix", "0")
    assert g.get_parameter("fft_radix") == "0"
