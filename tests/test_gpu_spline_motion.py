"""Splines that move far within a frame (a beating-heart-like phantom: every
scatterer swings up to 8 mm toward and away from a centre, neighbours together),
where the cull tests each line
against the box of the chunk's control points on the line's knot interval
(SplineSet::iboxes) rather than the box of all of them.  A box that missed a
point of its interval would drop that scatterer's echo from the line; the FP64
oracle, which evaluates every pair, would see the difference.  Every line is
compared, since a wrong cull drops whole lines' scatterers."""
import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def moving_splines(n=20000, lines=48, seed=81, degree=3, num_cp=12, phase_delay=False):
    wl = W.spline_mmode(n, lines, seed=seed, num_cp=num_cp, degree=degree, phase_delay=phase_delay)
    sp = wl.splines[0]
    cps = sp.control_points.astype(np.float64)
    # contraction toward a centre at 6 cm depth: each scatterer moves along the line to
    # the centre, by up to 8 mm, with the phase of the cycle (neighbours move together)
    centre = np.array([0.0, 0.0, 0.06])
    direction = centre - cps[:, :1, :]
    direction /= np.linalg.norm(direction, axis=2, keepdims=True) + 1e-12
    cycle = 0.5 - 0.5 * np.cos(np.linspace(0.0, 2.0 * np.pi, num_cp))[None, :, None]
    cps += 0.008 * direction * cycle
    wl.splines = [W.SplineSet(sp.degree, sp.knots, cps.astype(np.float32), sp.amplitudes)]
    return wl


@pytest.mark.parametrize("degree,phase", [(3, False), (2, True), (1, False)])
def test_moving_splines_match_oracle_on_every_line(degree, phase):
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = moving_splines(degree=degree, phase_delay=phase)
    g = RfSimulator("gpu")
    wl.apply(g)
    iq = g.simulate_frame()
    plan = g.get_parameter("launch_plan")
    g.close()
    assert "spline (interval boxes)" in plan, plan
    o = Oracle()
    wl.apply(o)
    ref = o.simulate_frame()
    errs = np.array([rel_l2(iq[i], ref[i]) for i in range(ref.shape[0])])
    print(f"[result] degree {degree} phase {phase}: frame {rel_l2(iq, ref):.2e}, worst line {errs.max():.2e}; "
          f"{plan.strip()!r}", flush=True)
    assert rel_l2(iq, ref) <= 1e-5
    assert errs.max() <= 1e-5


def test_static_splines_keep_chunk_boxes():
    """A phantom whose splines barely move (0.5 mm of jitter) gains nothing from the
    interval boxes, and the cull keeps the chunk boxes."""
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.spline_mmode(20000, 16, seed=82)
    g = RfSimulator("gpu")
    wl.apply(g)
    g.simulate_frame()
    plan = g.get_parameter("launch_plan")
    g.close()
    print(f"[result] static: {plan.strip()!r}", flush=True)
    assert "spline: " in plan and "interval boxes" not in plan
