"""The weight cutoff (parameter weight_cutoff): beam-profile weights below a
fraction c of the profile's peak magnitude count as zero and their pairs are
culled.  The rule is pinned against the FP64 oracle evaluating the same rule
(its weight_cutoff parameter): at c = 3e-3 the frame moves by ~c from the exact
sum, and the GPU must follow the oracle's rule to float rounding while the
distance to the exact sum stays ~c (a cull that drops too much or too little
fails one of the two).  The default c = 1e-8 stays ~1e-8 from the exact sum.

The launch's own counters (store_kernel_details: projection_weight_evals,
projection_lds_adds) are checked against a count of the same pairs in numpy."""
import math

import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "gauss_sector": lambda: W.fixed_sector_2d(20000, 16, seed=41),
    "gauss_sector_phase": lambda: W.fixed_sector_2d(20000, 16, seed=44, phase_delay=True),
    "gauss_linear_noarc": lambda: W.fixed_sector_2d(20000, 12, seed=45, arc=False),
    "lut_volume": lambda: W.volume_lut(30000, n_az=6, n_el=6, seed=42),
    "lut_volume_phase": lambda: W.volume_lut(30000, n_az=4, n_el=4, seed=46, phase_delay=True),
    "spline": lambda: W.spline_mmode(8000, 12, seed=43),
    "spline_equal_times": lambda: W.spline_mmode(8000, 12, seed=47, all_timestamps_equal=True),
}


def frames(wl, cutoff, mode="tiled"):
    from openbcsim_amd.pyrfsim import RfSimulator
    g = RfSimulator("gpu")
    wl.apply(g)
    g.set_parameter("weight_cutoff", str(cutoff))
    g.set_parameter("projection_mode", mode)
    iq = g.simulate_frame()
    g.close()
    o = Oracle()
    wl.apply(o)
    o.set_parameter("weight_cutoff", str(cutoff))
    same = o.simulate_frame()
    e = Oracle()
    wl.apply(e)
    return iq, same, e.simulate_frame()


@pytest.mark.parametrize("mode", ["tiled", "direct"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_cutoff_rule_matches_oracle(name, mode):
    if name == "gauss_linear_noarc":
        wl = CASES[name]()
        o, d, l, t, length = W.linear_scan(12, 0.03, wl.scan[4])
        wl.scan = (o, d, l, t, length)
        wl.params = dict(wl.params, use_arc_projection="off")
    else:
        wl = CASES[name]()
    print(f"[case] {name} {mode}", flush=True)
    iq, same, exact = frames(wl, 3e-3, mode)
    e_same, e_exact, ref = rel_l2(iq, same), rel_l2(iq, exact), rel_l2(same, exact)
    print(f"[result] {name} {mode} c=3e-3: vs same rule {e_same:.2e}, vs exact sum {e_exact:.2e} "
          f"(oracle's own distance {ref:.2e})", flush=True)
    assert e_same <= 1e-5
    assert abs(e_exact - ref) <= 0.05 * ref and ref > 1e-3


@pytest.mark.parametrize("name", ["gauss_sector", "lut_volume", "spline"])
def test_default_cutoff_is_invisible(name):
    wl = CASES[name]()
    iq, same, exact = frames(wl, 1e-8)
    print(f"[result] {name} default cutoff: vs exact sum {rel_l2(iq, exact):.2e}, "
          f"oracle rule vs exact {rel_l2(same, exact):.2e}", flush=True)
    assert rel_l2(same, exact) < 1e-7
    assert rel_l2(iq, exact) <= 1e-6


def numpy_pair_counts(wl, cutoff):
    """(weight evaluations, LDS additions) of a Gaussian fixed-point frame with arc
    projection: pairs whose bin lies in [0, num_rf) and whose exponent passes the
    cutoff, in the kernel's float32 arithmetic (left-to-right dot products)."""
    f32 = np.float32
    pts = wl.fixed[0].astype(np.float32)
    o, d, l, t, length = wl.scan
    e = np.stack([l[:, 1] * d[:, 2] - d[:, 1] * l[:, 2], d[:, 0] * l[:, 2] - l[:, 0] * d[:, 2],
                  l[:, 0] * d[:, 1] - d[:, 0] * l[:, 1]], 1).astype(np.float32)
    sl, se = wl.profile[1], wl.profile[2]
    a = f32(1.0 / (2.0 * float(f32(sl)) ** 2))
    b = f32(1.0 / (2.0 * float(f32(se)) ** 2))
    g_cut = f32(-math.log(float(f32(cutoff))))
    c = float(wl.params["sound_speed"])
    fs = wl.fs
    num_rf = wl.num_rf()
    evals = 0
    for k in range(len(d)):
        dx = pts[:, 0] - o[k, 0]
        dy = pts[:, 1] - o[k, 1]
        dz = pts[:, 2] - o[k, 2]
        r = dx * d[k, 0] + dy * d[k, 1] + dz * d[k, 2]
        ll = dx * l[k, 0] + dy * l[k, 1] + dz * l[k, 2]
        ee = dx * e[k, 0] + dy * e[k, 1] + dz * e[k, 2]
        r = np.copysign(np.sqrt(dx * dx + dy * dy + dz * dz), r)
        binv = np.floor(r.astype(np.float64) * 2.0 * fs / c + 0.5)
        x = ll * ll * a + ee * ee * b
        evals += int(np.count_nonzero((binv >= 0) & (binv < num_rf) & (x <= g_cut)))
    return evals


def test_launch_counters():
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.fixed_sector_2d(20000, 16, seed=48)
    counts = {}
    for cutoff in (1e-8, 3e-3):
        g = RfSimulator("gpu")
        wl.apply(g)
        g.set_parameter("weight_cutoff", str(cutoff))
        g.set_parameter("projection_counters", "on")
        g.simulate_frame()
        ev = g.get_debug_data("projection_weight_evals")
        ad = g.get_debug_data("projection_lds_adds")
        pairs = g.get_debug_data("projection_chunk_line_pairs")
        assert len(pairs) == 1 and 64 * pairs[0] >= ev[0]  # each kept (chunk, line) pair is 64 lanes
        g.close()
        assert len(ev) == 1 and len(ad) == 1
        counts[cutoff] = (ev[0], ad[0])
        want = numpy_pair_counts(wl, cutoff)
        print(f"[result] cutoff {cutoff}: evals {ev[0]:.0f} (numpy {want}), adds {ad[0]:.0f}", flush=True)
        assert abs(ev[0] - want) <= 1e-5 * want + 2
        assert ad[0] <= ev[0]  # real bins: one addition per non-zero weight
        assert ad[0] >= 0.999 * ev[0]  # amplitudes N(0, 1): zero only by underflow
    assert counts[3e-3][0] < 0.7 * counts[1e-8][0]
