"""Parity of the HIP path (through the C-ABI) with the FP64 oracle and with the
reference's own CPU algorithm.

Tolerance: relative L2 error of the complex IQ
frame <= 1e-4 against the FP64 oracle, because float atomics make the summation
order nondeterministic.  The projection bins are checked on their own too:
every bin that carries signal is the bin the oracle chose (indices are integer
work and must be exact), values within 1e-5.

The product drops profile weights below 1e-8 of the peak by default (the
weight_cutoff parameter); bins are compared with the oracle under the same
rule, and every IQ frame also with the oracle's exact sum (no cutoff), the
reference's semantics, at the same 1e-4 bar.
"""
import os

import numpy as np
import pytest

from backends import Oracle, rel_l2
from golden.make_golden import golden_cases, input_digest
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-4


def gpu_sim():
    from openbcsim_amd.pyrfsim import RfSimulator
    return RfSimulator("gpu")


def run_pair(wl):
    """The GPU frame and bins, and an oracle under the GPU's weight cutoff."""
    print(f"[case] {wl.name} lines={wl.num_lines} scatterers={wl.num_scatterers} "
          f"N_fft={wl.fft_len()}", flush=True)
    g = gpu_sim()
    wl.apply(g)
    iq = g.simulate_frame()
    bins = g.get_projection()
    o = Oracle()
    wl.apply(o)
    o.set_parameter("weight_cutoff", g.get_parameter("weight_cutoff"))
    return g, iq, bins, o


def exact_oracle_err(wl, iq):
    """Relative L2 error against the oracle's sum over every pair (no cutoff)."""
    o = Oracle()
    wl.apply(o)
    return rel_l2(iq, o.simulate_frame())


def check_bins(bins, ob, tol=1e-5):
    strong = np.maximum(np.abs(bins), np.abs(ob)) > 1e-20 * max(1.0, float(np.abs(ob).max()))
    assert np.all(np.abs(ob[strong]) > 0) and np.all(np.abs(bins[strong]) > 0), "bin index mismatch"
    err = rel_l2(bins, ob)
    assert err < tol, err


CASES = {
    "fixed_gauss_arc": lambda: W.fixed_sector_2d(20000, 16, seed=21),
    "fixed_gauss_arc_phase": lambda: W.fixed_sector_2d(20000, 16, seed=22, phase_delay=True),
    "fixed_gauss_noarc_dec3": lambda: W.fixed_sector_2d(20000, 12, seed=23, arc=False, decimation=3),
    "spline_cubic_dec4": lambda: W.spline_mmode(10000, 16, seed=24),
    "spline_cubic_phase": lambda: W.spline_mmode(10000, 16, seed=25, phase_delay=True, decimation=2),
    "spline_equal_times": lambda: W.spline_mmode(10000, 16, seed=26, all_timestamps_equal=True),
    "spline_degree1": lambda: W.spline_mmode(5000, 8, seed=27, num_cp=5, degree=1),
    "spline_degree4": lambda: W.spline_mmode(5000, 8, seed=28, num_cp=12, degree=4),
    "lut_volume_dec2": lambda: W.volume_lut(50000, n_az=6, n_el=6, seed=29),
    "lut_volume_phase": lambda: W.volume_lut(50000, n_az=4, n_el=4, seed=30, phase_delay=True, decimation=1),
    # decimations whose largest power-of-two factor folds the inverse transform
    "fixed_gauss_dec1": lambda: W.fixed_sector_2d(20000, 12, seed=31, decimation=1),
    "fixed_gauss_dec8_short": lambda: W.fixed_sector_2d(8000, 12, seed=32, line_length=0.028, decimation=8),
    "fixed_gauss_phase_dec6": lambda: W.fixed_sector_2d(20000, 12, seed=33, phase_delay=True, decimation=6),
    "fixed_gauss_dec16_tiny": lambda: W.fixed_sector_2d(3000, 8, seed=34, line_length=0.002, decimation=16),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_iq_matches_oracle(name):
    wl = CASES[name]()
    wl.name = name
    g, iq, bins, o = run_pair(wl)
    check_bins(bins, o.project())
    err = rel_l2(iq, o.simulate_frame())
    exact = exact_oracle_err(wl, iq)
    print(f"[result] {name} rel_l2={err:.3e} (exact sum: {exact:.3e})", flush=True)
    assert err <= TOL and exact <= TOL


def test_mixed_datasets_linear_scan():
    mix = W.fixed_sector_2d(4000, 10, seed=31, line_length=0.06, phase_delay=True)
    o_, d, l, t, length = W.linear_scan(10, 0.03, 0.06)
    spl = W.spline_mmode(3000, 10, seed=32, line_length=0.06)
    mix.scan = (o_, d, l, spl.scan[3], length)
    mix.fixed.append(W.fixed_sector_2d(2000, 10, seed=33).fixed[0])
    mix.splines = spl.splines + W.spline_mmode(1000, 10, seed=34, num_cp=7, degree=2).splines
    mix.name = "mixed_linear"
    g, iq, bins, o = run_pair(mix)
    check_bins(bins, o.project())
    assert rel_l2(iq, o.simulate_frame()) <= TOL
    assert exact_oracle_err(mix, iq) <= TOL


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_iq_matches_reference_golden(name):
    """The reference CPU algorithm's own frames (compiled from its sources).  The
    reference is itself ~5e-5 from an exact evaluation (float demodulation phase,
    recursive-FFT twiddle recurrence), so the bar is the same 1e-4."""
    wl = golden_cases()[name]
    fx = np.load(os.path.join(HERE, "golden", f"{name}.npz"))
    assert str(fx["digest"]) == input_digest(wl)
    g = gpu_sim()
    wl.apply(g)
    err_cut = rel_l2(g.simulate_frame(), fx["iq"])  # the default weight cutoff
    g.set_parameter("weight_cutoff", "0")  # the reference keeps every pair: so do the bins here
    iq = g.simulate_frame()
    err = rel_l2(iq, fx["iq"])
    print(f"[result] golden {name} rel_l2={err:.3e} (default weight cutoff: {err_cut:.3e})", flush=True)
    assert err <= TOL and err_cut <= TOL
    bins = g.get_projection().ravel()
    ref = np.zeros_like(bins)
    ref[fx["bin_index"]] = fx["bin_value"]
    # the same bins are non-zero: every bin above 1e-9 of the largest (the
    # fixed-point bins resolve contributions down to ~1e-15 of it; smaller sums
    # are zero there while the reference's float tails reach 1e-40)
    strong = np.maximum(np.abs(bins), np.abs(ref)) > 1e-9 * np.abs(ref).max()
    assert np.all(np.abs(ref[strong]) > 0) and np.all(np.abs(bins[strong]) > 0)


def test_full_size_volume_sampled_lines():
    """Full 1M-scatterer, 64x64-line volume with the LUT profile; the oracle
    re-evaluates a sample of 12 lines against all 1M scatterers."""
    wl = W.volume_lut()
    g = gpu_sim()
    wl.apply(g)
    iq = g.simulate_frame()
    assert iq.shape == (4096, 4870)
    pick = np.array([0, 63, 1000, 2016, 2047, 2048, 2080, 3000, 3500, 4032, 4095, 1234])
    sub = W.volume_lut()
    o_, d, l, t, length = sub.scan
    sub.scan = (o_[pick], d[pick], l[pick], t[pick], length)
    o = Oracle()
    sub.apply(o)
    err = rel_l2(iq[pick], o.simulate_frame())
    print(f"[result] full volume sampled rel_l2={err:.3e}", flush=True)
    assert err <= TOL


def test_full_size_spline_sampled_lines_and_linearity():
    wl = W.spline_mmode()
    g = gpu_sim()
    wl.apply(g)
    iq = g.simulate_frame()
    assert iq.shape == (256, 1948)
    pick = np.array([0, 17, 128, 200, 255])
    sub = W.spline_mmode()
    o_, d, l, t, length = sub.scan
    sub.scan = (o_[pick], d[pick], l[pick], t[pick], length)
    o = Oracle()
    sub.apply(o)
    assert rel_l2(iq[pick], o.simulate_frame()) <= TOL
    # amplitude x2 doubles the frame (power-of-two scaling is exact up to summation order)
    wl2 = W.spline_mmode()
    wl2.splines[0].amplitudes = wl2.splines[0].amplitudes * np.float32(2.0)
    g2 = gpu_sim()
    wl2.apply(g2)
    assert rel_l2(g2.simulate_frame(), 2.0 * iq) < 1e-6


def test_decimation_is_subsampling_and_repeatable():
    wl = W.fixed_sector_2d(30000, 8, seed=41, phase_delay=True)
    g = gpu_sim()
    wl.apply(g)
    a = g.simulate_frame()
    b = g.simulate_frame()
    assert rel_l2(b, a) < 1e-6
    g.set_parameter("radial_decimation", "5")
    c = g.simulate_frame()
    assert c.shape == (8, (7792 + 4) // 5)
    assert rel_l2(c, a[:, ::5]) < 1e-5
    # pyrfsim layout (num_samples, num_lines)
    assert g.simulate_lines().shape == (c.shape[1], c.shape[0])


def test_edge_cases():
    # one scatterer on one line, short line (small FFT), scatterers behind the probe and beyond
    wl = W.fixed_sector_2d(1, 1, line_length=0.002)
    wl.fixed = [np.array([[0, 0, 0.001, 1.0], [0, 0, -0.001, 1.0], [0, 0, 0.5, 1.0]], np.float32)]
    g, iq, bins, o = run_pair(wl)
    assert rel_l2(iq, o.simulate_frame()) <= TOL
    # empty phantom: the GPU path returns a zero frame (GpuAlgorithm.cpp has no scatterer check)
    g2 = gpu_sim()
    wl2 = W.fixed_sector_2d(10, 4, line_length=0.01)
    wl2.fixed = [np.zeros((0, 4), np.float32)]
    wl2.apply(g2)
    assert np.all(g2.simulate_frame() == 0)
    # decimation larger than the line
    g2.set_parameter("radial_decimation", "100000")
    assert g2.simulate_frame().shape == (4, 1)


def test_errors_through_the_abi():
    from openbcsim_amd.pyrfsim import BcsimError, RfSimulator
    with pytest.raises(BcsimError, match="Illegal algorithm type"):
        RfSimulator("cpu")
    g = gpu_sim()
    with pytest.raises(BcsimError, match="Excitation must be configured before scan sequence"):
        g.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.05, np.array([[1.0, 0, 0]]), np.zeros(1))
    exc, c = W.make_gaussian_excitation()
    g.set_excitation(exc, c, W.FS, W.FC)
    with pytest.raises(BcsimError, match="Not orthogonal"):
        g.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.05, np.array([[1.0, 0.2, 0]]), np.zeros(1))
    g.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.05, np.array([[1.0, 0, 0]]), np.zeros(1))
    with pytest.raises(BcsimError, match="No beam profile"):
        g.simulate_frame()
    with pytest.raises(BcsimError, match="maximum spline degree"):
        g.add_spline_scatterers(5, np.zeros(12), np.zeros((1, 6, 3)), np.zeros(1))
    with pytest.raises(BcsimError, match="illegal parameter name"):
        g.set_parameter("num_cpu_cores", "4")
    with pytest.raises(BcsimError, match="cannot change CUDA device"):
        g.set_parameter("gpu_device", "0")
    with pytest.raises(BcsimError, match="Too few radial samples"):
        g.set_lut_beam_profile(0, 1, 0, 1, 0, 1, np.zeros((1, 4, 4)))
    g.set_analytical_beam_profile(1e-3, 1e-3)
    # 0.3 m gives 19481 RF samples: FFT length 32768 exceeds the 16384-point LDS transform
    g.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.3, np.array([[1.0, 0, 0]]), np.zeros(1))
    with pytest.raises(BcsimError, match="Too many RF samples"):
        g.simulate_frame()
    # a timestamp outside the knot span (bspline.hpp:81)
    g.set_scan_sequence(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), 0.05, np.array([[1.0, 0, 0]]), np.ones(1))
    sp = W.spline_mmode(10, 1).splines[0]
    g.add_spline_scatterers(sp.degree, sp.knots, sp.control_points, sp.amplitudes)
    with pytest.raises(BcsimError, match="knot interval"):
        g.simulate_frame()
    assert g.get_total_num_scatterers() == 10
    assert "gfx950" in g.get_parameter("cur_device_name")


def test_kernel_details_debug_data():
    wl = W.fixed_sector_2d(5000, 4, seed=5)
    g = gpu_sim()
    g.set_parameter("store_kernel_details", "on")
    wl.apply(g)
    g.simulate_frame()
    for key in ("fixed_projection_kernel_ms", "kernel_convolve_demodulate_ms", "kernel_memcpy_ms"):
        v = g.get_debug_data(key)
        assert len(v) >= 1 and all(x >= 0 for x in v)
    from openbcsim_amd.pyrfsim import BcsimError
    with pytest.raises(BcsimError, match="invalid debug data key"):
        g.get_debug_data("kernel_projection_ms")


@pytest.mark.parametrize("name", ["lut_volume_dec2", "fixed_gauss_arc_phase", "spline_cubic_dec4",
                                  "spline_equal_times", "fixed_gauss_noarc_dec3"])
def test_tiled_matches_direct_projection(name):
    """The culled tile kernel skips only pairs whose contribution is exactly zero:
    its bins equal the direct kernel's (every pair evaluated) up to summation order."""
    wl = CASES[name]()
    out = {}
    for mode in ("direct", "tiled"):
        g = gpu_sim()
        g.set_parameter("projection_mode", mode)
        wl.apply(g)
        iq = g.simulate_frame()
        out[mode] = (iq, g.get_projection())
    assert np.array_equal(np.abs(out["direct"][1]) > 0, np.abs(out["tiled"][1]) > 0)
    assert rel_l2(out["tiled"][1], out["direct"][1]) < 1e-6
    assert rel_l2(out["tiled"][0], out["direct"][0]) < 1e-6


@pytest.mark.parametrize("phase", [False, True])
def test_tiled_matches_direct_full_volume(phase):
    """The tiled kernel against the direct one (every pair, no culling, no chunk
    list) on the full volume, whose tiles (4 lines of real bins, 2 of complex
    ones) hold the chunk list as one-word entries (chunk << 4 | line mask)."""
    wl = W.volume_lut(phase_delay=phase)
    res, plan = {}, ""
    for mode in ("direct", "tiled"):
        g = gpu_sim()
        g.set_parameter("projection_mode", mode)
        wl.apply(g)
        res[mode] = g.simulate_frame()
        if mode == "tiled":
            plan = g.get_parameter("launch_plan")
        g.close()
    err = rel_l2(res["tiled"], res["direct"])
    print(f"[result] full volume (phase delay {phase}) tiled vs direct rel_l2={err:.3e}; plan {plan.strip()!r}",
          flush=True)
    assert ("lines=2" if phase else "lines=4") in plan and "packed" in plan
    assert err < 1e-6


def test_scan_length_changes_between_frames():
    """One simulator whose scan length, and so its FFT length (2048 - 16384) and
    output size, changes between frames, with the output shape queried before each
    frame (simulate_frame does): every frame equals a fresh simulator's and the
    oracle's.  The convolution tables must follow the new frame shape."""
    from openbcsim_amd.pyrfsim import RfSimulator
    g = RfSimulator("gpu")
    first = True
    for length in (0.028, 0.12, 0.15, 0.06, 0.15, 0.028, 0.12):
        wl = W.fixed_sector_2d(6000, 12, seed=91, line_length=length, decimation=2)
        if first:
            wl.apply(g)
            first = False
        else:
            o, d, l, t, ln = wl.scan
            g.set_scan_sequence(o, d, ln, l, t, False)
        iq = g.simulate_frame()
        fresh = gpu_sim()
        wl.apply(fresh)
        ref = fresh.simulate_frame()
        fresh.close()
        e = rel_l2(iq, ref)
        print(f"[result] line length {length} N_fft={wl.fft_len()} vs fresh simulator {e:.2e}", flush=True)
        assert iq.shape == ref.shape and e < 1e-6
    o = Oracle()
    wl.apply(o)
    assert rel_l2(iq, o.simulate_frame()) <= TOL


def test_tile_order_keeps_frames():
    """Blocks of large tile launches run heaviest first, ordered by the previous
    frame's counts, and visit the chunks near to far from the scan origin: the
    frames must not change beyond float summation order, from the first frame
    (no counts yet) on, when the number of blocks changes between frames, and
    with several datasets (one order per dataset)."""
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.volume_lut(300_000, n_az=64, n_el=40, seed=41)
    wl.fixed.append(W.volume_lut(200_000, n_az=2, n_el=2, seed=42).fixed[0])
    sims = {}
    for mode in ("off", "on"):
        s = RfSimulator("gpu")
        wl.apply(s)
        for key in ("tile_order", "chunk_order"):
            s.set_parameter(key, mode)
            assert s.get_parameter(key) == mode
        sims[mode] = s
    o_, d, l, t, length = wl.scan
    for frame, keep in enumerate([2560, 2560, 2560, 2048, 2560, 640]):
        for s in sims.values():
            s.set_scan_sequence(o_[:keep], d[:keep], length, l[:keep], t[:keep], False)
        ref = sims["off"].simulate_frame()
        iq = sims["on"].simulate_frame()
        bins_on, bins_off = sims["on"].get_projection(), sims["off"].get_projection()
        e = rel_l2(iq, ref)
        print(f"[result] frame {frame} lines {keep}: order on vs off rel_l2={e:.2e}", flush=True)
        assert iq.shape == ref.shape and e < 1e-6
        assert np.array_equal(np.abs(bins_on) > 0, np.abs(bins_off) > 0)
    sub = W.volume_lut(300_000, n_az=64, n_el=40, seed=41)
    sub.fixed.append(wl.fixed[1])
    pick = np.array([0, 5, 300, 639])
    sub.scan = (o_[pick], d[pick], l[pick], t[pick], length)
    o = Oracle()
    sub.apply(o)
    err = rel_l2(iq[pick], o.simulate_frame())
    print(f"[result] ordered frame vs oracle (4 lines) rel_l2={err:.2e}", flush=True)
    assert err <= TOL


def test_tile_order_many_blocks_and_windows():
    """Block ordering over thousands of blocks (one line per tile) and over
    (tile, window) blocks: frames equal those of the index order."""
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.volume_lut(200_000, n_az=64, n_el=64, seed=44)
    res = {}
    for mode in ("off", "on"):
        s = RfSimulator("gpu")
        wl.apply(s)
        s.set_parameter("tile_order", mode)
        out = []
        for lines, windows in ((1, 1), (2, 2), (4, 1)):
            s.set_parameter("tile_lines", str(lines))
            s.set_parameter("tile_windows", str(windows))
            out.append([s.simulate_frame() for _ in range(2)])
        res[mode] = out
        s.close()
    for k, (lines, windows) in enumerate(((1, 1), (2, 2), (4, 1))):
        for f in range(2):
            e = rel_l2(res["on"][k][f], res["off"][k][f])
            print(f"[result] tile_lines={lines} windows={windows} frame {f}: order on vs off {e:.2e}", flush=True)
            assert e < 1e-6


@pytest.mark.parametrize("phase", [False, True])
def test_smallest_transform_every_decimation(phase):
    """A 5-sample excitation and a 0.18 mm line give the smallest transform
    (16 points): the real-input path then runs an 8-point forward transform in
    one pass, and no decimation can fold the inverse below 16 points."""
    rng = np.random.default_rng(45)
    for dec in (1, 2, 3, 4, 8, 16):
        wl = W.fixed_sector_2d(1, 6, line_length=0.00018, decimation=dec, phase_delay=phase)
        wl.excitation = np.array([0.2, -0.7, 1.0, -0.6, 0.25], np.float32)
        wl.center_index = 2
        pts = np.zeros((400, 4), np.float32)
        pts[:, 0] = rng.uniform(-0.0005, 0.0005, 400)
        pts[:, 2] = rng.uniform(0.00002, 0.00017, 400)
        pts[:, 3] = rng.standard_normal(400)
        wl.fixed = [pts]
        assert wl.fft_len() == 16, wl.fft_len()
        g, iq, bins, o = run_pair(wl)
        err = rel_l2(iq, o.simulate_frame())
        print(f"[result] N_fft=16 decimation={dec} phase={phase}: {iq.shape} rel_l2={err:.2e}", flush=True)
        assert err <= TOL


def test_host_output_paths_agree():
    """simulate_lines into pageable host memory (pinned staging, piece by piece)
    and into device memory give the same frame (up to float summation order),
    above and below the staging threshold and across several staging pieces."""
    from openbcsim_amd.hip_rt import DeviceBuffer
    for wl in (W.volume_lut(100_000, n_az=64, n_el=64, seed=46),   # 160 MB: five 32 MB pieces
               W.fixed_sector_2d(5000, 8, seed=47)):                # 0.5 MB: direct copy
        s = gpu_sim()
        wl.apply(s)
        s.set_parameter("tile_order", "off")  # same block order in both frames
        host = s.simulate_frame()
        L, S = s.output_shape()
        dev = DeviceBuffer(L * S * 8)
        s.simulate_lines_device(dev.address, L * S)
        s.synchronize()
        ref = np.empty((L, S), np.complex64)
        dev.to_host(ref)
        d = float(np.abs(host - ref).max()) / max(float(np.abs(ref).max()), 1e-30)
        print(f"[result] {wl.name} {L}x{S} ({host.nbytes / 1e6:.1f} MB): host vs device frame max rel diff {d:.2e}",
              flush=True)
        assert np.allclose(host, ref, rtol=1e-6, atol=1e-6 * float(np.abs(ref).max()))
        s.close()


def test_phantom_files_through_the_gpu_path(tmp_path):
    """A fixed phantom, a LUT profile and a scan loaded from HDF5 files of the
    reference's formats (openbcsim_amd.hdf) into the GPU simulator: the frame of
    the same arrays in the FP64 oracle, at the parity bar."""
    from openbcsim_amd import hdf
    vol = W.volume_lut(20000, n_az=8, n_el=8, seed=9, lut_shape=(64, 16, 16))
    (r0, r1, l0, l1, e0, e1), table = vol.profile[1], vol.profile[2]
    o, d, l, t, length = vol.scan
    paths = {k: str(tmp_path / f"{k}.h5") for k in ("fixed", "lut", "scan")}
    hdf.save_fixed_scatterers(paths["fixed"], vol.fixed[0])
    hdf.save_beam_profile(paths["lut"], table, (r0, r1), (l0, l1), (e0, e1))
    hdf.save_scan_sequence(paths["scan"], o, d, l, t, length)
    g = gpu_sim()
    for k, v in vol.params.items():
        g.set_parameter(k, str(v))
    g.set_excitation(vol.excitation, vol.center_index, vol.fs, vol.demod_freq)
    hdf.set_scan_sequence_from_hdf(g, paths["scan"])
    hdf.set_fixed_scatterers_from_hdf(g, paths["fixed"])
    hdf.set_beam_profile_from_hdf(g, paths["lut"])
    iq = g.simulate_frame()
    g.close()
    o_ = Oracle()
    for k, v in vol.params.items():
        o_.set_parameter(k, str(v))
    o_.set_excitation(vol.excitation, vol.center_index, vol.fs, vol.demod_freq)
    o_.set_scan_sequence(o, d, length, l, t)
    o_.add_fixed_scatterers(vol.fixed[0])
    o_.set_lut_beam_profile(r0, r1, l0, l1, e0, e1, table / table.max())
    ref = o_.simulate_frame()
    e = rel_l2(iq, ref)
    print(f"[result] phantom files through the GPU path: rel L2 vs oracle {e:.2e}", flush=True)
    assert np.abs(ref).max() > 0 and e <= 1e-4


@pytest.mark.parametrize("phase", [False, True])
def test_lut_unequal_lateral_and_elevation(phase):
    """A LUT of unequal lateral and elevational sizes and extents with random
    samples (no symmetry between l and e): a swapped lateral/elevational index
    or extent, or a transposed table, moves the frame far from the oracle's."""
    import dataclasses
    rng = np.random.default_rng(21)
    vol = W.volume_lut(20000, n_az=6, n_el=5, seed=22, line_length=0.06, lut_shape=(40, 12, 20),
                       phase_delay=phase)
    table = rng.uniform(0.0, 1.0, (40, 12, 20)).astype(np.float32)
    wl = dataclasses.replace(vol, profile=("lut", (0.0, 0.06, -0.012, 0.009, -0.006, 0.014), table))
    o_ = Oracle()
    wl.apply(o_)
    ref = o_.simulate_frame()
    for mode in ("tiled", "direct"):
        g = gpu_sim()
        g.set_parameter("projection_mode", mode)
        wl.apply(g)
        iq = g.simulate_frame()
        g.close()
        e = rel_l2(iq, ref)
        print(f"[result] unequal LUT phase={phase} {mode}: rel L2 vs oracle {e:.2e}", flush=True)
        assert np.abs(ref).max() > 0 and e <= 1e-4
    # the same table transposed in l and e (extents swapped) is a different frame
    wt = dataclasses.replace(vol, profile=("lut", (0.0, 0.06, -0.006, 0.014, -0.012, 0.009),
                                            np.ascontiguousarray(table.transpose(0, 2, 1))))
    o_t = Oracle()
    wt.apply(o_t)
    assert rel_l2(o_t.simulate_frame(), ref) > 1e-2


@pytest.mark.parametrize("arc", [True, False])
def test_linear_scan_with_lut(arc):
    """Lines that do not share an origin (a linear scan: the per-line cull and
    projection of the tile kernel) with a LUT profile, arc projection on and off."""
    import dataclasses
    rng = np.random.default_rng(23)
    vol = W.volume_lut(20000, n_az=2, n_el=2, seed=24, line_length=0.06, lut_shape=(40, 10, 14))
    o, d, l, t, length = W.linear_scan(24, 0.04, 0.06)
    table = rng.uniform(0.0, 1.0, (40, 10, 14)).astype(np.float32)
    params = dict(vol.params, use_arc_projection="on" if arc else "off")
    wl = dataclasses.replace(vol, scan=(o, d, l, t, length), params=params,
                             profile=("lut", (0.0, 0.06, -0.01, 0.01, -0.007, 0.007), table))
    o_ = Oracle()
    wl.apply(o_)
    ref = o_.simulate_frame()
    g = gpu_sim()
    wl.apply(g)
    iq = g.simulate_frame()
    g.close()
    e = rel_l2(iq, ref)
    print(f"[result] linear scan with LUT arc={arc}: rel L2 vs oracle {e:.2e}", flush=True)
    assert np.abs(ref).max() > 0 and e <= 1e-4


def test_radial_windows_visit_every_reachable_chunk():
    """Radial windows visit only the chunks whose boxes can reach their distances
    (bounds kept with each common-origin chunk order).  A phantom of three dense
    clusters at different depths plus a sparse background has chunks whose boxes
    span the gaps where the z-curve jumps; frames in 2 and 4 windows must equal the
    single-window frame, and the oracle on sampled lines."""
    from openbcsim_amd.pyrfsim import RfSimulator
    rng = np.random.default_rng(47)
    parts = []
    for centre, half, count in (((0.0, 0.0, 0.02), 0.004, 60000), ((0.01, -0.005, 0.07), 0.006, 60000),
                                ((-0.008, 0.006, 0.12), 0.005, 60000), ((0.0, 0.0, 0.075), 0.03, 20000)):
        p = np.empty((count, 4), np.float32)
        p[:, :3] = rng.uniform(-half, half, (count, 3)) + np.array(centre)
        p[:, 3] = rng.standard_normal(count)
        parts.append(p)
    wl = W.volume_lut(1000, n_az=32, n_el=32, seed=47)
    wl.fixed = [np.concatenate(parts)]
    frames, plans = {}, {}
    for windows in (1, 2, 4):
        g = RfSimulator("gpu")
        wl.apply(g)
        g.set_parameter("tile_windows", str(windows))
        g.set_parameter("tile_lines", "4")
        frames[windows] = g.simulate_frame()
        plans[windows] = g.get_parameter("launch_plan").splitlines()[0]
        g.close()
    for windows in (2, 4):
        e = rel_l2(frames[windows], frames[1])
        print(f"[result] {plans[windows]!r}: vs one window rel_l2={e:.2e}", flush=True)
        assert f"windows={windows}" in plans[windows]
        assert e < 1e-6
    pick = np.array([0, 200, 527, 1023])
    o_, d, l, t, length = wl.scan
    sub = W.volume_lut(1000, n_az=32, n_el=32, seed=47)
    sub.fixed = wl.fixed
    sub.scan = (o_[pick], d[pick], l[pick], t[pick], length)
    o = Oracle()
    sub.apply(o)
    o.set_parameter("weight_cutoff", "1e-8")
    err = rel_l2(frames[4][pick], o.simulate_frame())
    print(f"[result] 4 windows vs oracle (4 lines) rel_l2={err:.2e}", flush=True)
    assert err <= TOL
