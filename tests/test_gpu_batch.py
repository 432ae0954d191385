"""Batched frames (persistent launches, HIP graph replay) against single-simulator
frames and the FP64 oracle."""
import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_sim(wl):
    from openbcsim_amd.pyrfsim import RfSimulator
    s = RfSimulator("gpu")
    wl.apply(s)
    return s


def mixed_jobs():
    """Uneven jobs: fixed and spline, analytic and LUT, arc off, phase delay, several
    FFT lengths and decimations, two datasets in one job."""
    jobs = [
        W.fixed_sector_2d(20000, 40, seed=61, line_length=0.06, decimation=2),
        W.spline_mmode(8000, 24, seed=62, line_length=0.028, decimation=1, phase_delay=True),
        W.volume_lut(30000, n_az=5, n_el=4, seed=63, line_length=0.12, phase_delay=True),
        W.fixed_sector_2d(5000, 7, seed=64, line_length=0.15, arc=False, decimation=3),
        W.spline_mmode(3000, 33, seed=65, num_cp=6, degree=2, all_timestamps_equal=True),
    ]
    two = W.fixed_sector_2d(4000, 16, seed=66, line_length=0.04)
    two.fixed.append(W.fixed_sector_2d(3000, 16, seed=67).fixed[0])
    two.splines = W.spline_mmode(2000, 16, seed=68).splines
    two.scan = two.scan[:3] + (W.spline_mmode(2000, 16, seed=68).scan[3], two.scan[4])
    jobs.append(two)
    return jobs


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_single_and_oracle(use_graph):
    from openbcsim_amd.batch import Batch
    jobs = mixed_jobs()
    sims = [make_sim(wl) for wl in jobs]
    singles = [s.simulate_frame() for s in sims]
    b = Batch(sims, use_graph=use_graph)
    frames = b.simulate()
    assert len(frames) == len(jobs)
    for i, (wl, f, s) in enumerate(zip(jobs, frames, singles)):
        assert f.shape == s.shape, i
        err_s = rel_l2(f, s)
        o = Oracle()
        wl.apply(o)
        err_o = rel_l2(f, o.simulate_frame())
        print(f"[result] batch job {i} vs single {err_s:.2e} vs oracle {err_o:.2e}", flush=True)
        assert err_s < 1e-6 and err_o <= 1e-4


def test_graph_replay_and_reconfiguration():
    """Replays of the captured graph, with eager frames of the same handles in
    between, all agree; changing a job's parameter rebuilds the plan."""
    from openbcsim_amd.batch import Batch
    from openbcsim_amd.hip_rt import DeviceBuffer
    jobs = mixed_jobs()[:4]
    sims = [make_sim(wl) for wl in jobs]
    b = Batch(sims)
    total, layout = b.layout()
    out = DeviceBuffer(total * 8)
    ref = None
    for rep in range(4):
        b.simulate_device(out.address, total)
        b.synchronize()
        host = np.empty(total, np.complex64)
        out.to_host(host)
        if ref is None:
            ref = host.copy()
        else:
            assert rel_l2(host, ref) < 1e-6, rep
        sims[rep % len(sims)].simulate_frame()  # eager work between replays
    # reconfigure job 0: decimation 2 -> 4 changes its layout
    sims[0].set_parameter("radial_decimation", "4")
    total2, layout2 = b.layout()
    assert layout[0][2] == (3896 + 1) // 2 and layout2[0][2] == (3896 + 3) // 4
    frames = b.simulate()
    assert rel_l2(frames[0], sims[0].simulate_frame()) < 1e-6


def test_batch_config3_jobs_match_single():
    """The 64 uneven jobs of the batched workload at a tenth of their scatterers:
    every job's batched frame equals its single frame."""
    from openbcsim_amd.batch import Batch
    jobs = W.batched_jobs(64, seed=3, scale=0.1)
    sims = [make_sim(wl) for wl in jobs]
    b = Batch(sims)
    frames = b.simulate()
    worst = 0.0
    for i, (s, f) in enumerate(zip(sims, frames)):
        e = rel_l2(f, s.simulate_frame())
        worst = max(worst, e)
    print(f"[result] 64-job batch worst rel_l2 vs single = {worst:.2e}", flush=True)
    assert worst < 1e-6


def test_full_scale_batch_jobs_against_oracle():
    """The 64 uneven jobs of the batched workload at full scale (10k-2M
    scatterers, fixed and spline, analytic and LUT, FFT 2048-16384): eight of
    them, from the smallest to the largest, taken from one batched frame and
    compared on four lines each with the FP64 oracle's exact sum."""
    from openbcsim_amd.batch import Batch
    jobs = W.batched_jobs(64, seed=3)
    sims = [make_sim(wl) for wl in jobs]
    b = Batch(sims)
    frames = b.simulate()
    order = sorted(range(len(jobs)), key=lambda j: jobs[j].num_scatterers)
    picks = [order[k] for k in (0, 9, 18, 27, 36, 45, 54, 63)]
    kinds = set()
    worst = 0.0
    for j in picks:
        wl = jobs[j]
        kinds.add(("spline" if wl.splines else "fixed", wl.profile[0]))
        o, d, l, t, length = wl.scan
        sel = np.unique(np.linspace(0, o.shape[0] - 1, 4).astype(int))
        wl.scan = (o[sel], d[sel], l[sel], t[sel], length)
        orc = Oracle()
        wl.apply(orc)
        e = rel_l2(frames[j][sel], orc.simulate_frame())
        print(f"[result] job {j}: {wl.num_scatterers} scatterers, N_fft {wl.fft_len()}, "
              f"{'spline' if wl.splines else 'fixed'}/{wl.profile[0]}: rel_l2 vs oracle {e:.2e}", flush=True)
        worst = max(worst, e)
    assert worst <= 1e-4
    assert len(kinds) >= 3, kinds  # the picks span the job kinds
    b.close()
    for s in sims:
        s.close()
