"""Additive noise on the time-projection samples (CpuAlgorithm.cpp:340-347: complex
noise, each component N(0, noise_amplitude^2), added to samples [0, num_rf) before
the convolution).  The reference draws from a shared generator, so parity is
statistical: the IQ of noise alone has the variance the filter fixes,
E|IQ|^2 = 2 a^2 sum |h_a|^2 with h_a the analytic excitation the convolution
applies (numpy), and Gaussian marginals.  The counter-based stream of this build
is also checked for determinism per seed, advance per frame and linearity."""
import numpy as np
import pytest

from backends import rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def noise_sim(amplitude, seed, scatterers=0, decimation=2, lines=64):
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.fixed_sector_2d(max(scatterers, 1), lines, seed=5, line_length=0.06, decimation=decimation)
    if scatterers == 0:
        wl.fixed = [wl.fixed[0][:0]]  # no scatterers: the frame is filtered noise
    s = RfSimulator("gpu")
    wl.apply(s)
    s.set_parameter("noise_amplitude", str(amplitude))
    s.set_parameter("noise_seed", str(seed))
    return s, wl


def analytic_energy(wl):
    """sum |h_a|^2, h_a = IFFT_N(mask * FFT_N(excitation)) (BeamConvolver.cpp:77-91)."""
    N = wl.fft_len()
    X = np.fft.fft(wl.excitation.astype(np.float64), N)
    mask = np.zeros(N)
    mask[0] = 1.0
    mask[1:N // 2] = 2.0
    mask[N // 2] = 1.0
    h = np.fft.ifft(mask * X)
    return float(np.sum(np.abs(h) ** 2))


def test_noise_variance_and_gaussian_marginals():
    a = 0.7
    s, wl = noise_sim(a, seed=1234)
    frames = [s.simulate_frame() for _ in range(4)]
    iq = np.concatenate([f[:, 60:-60].ravel() for f in frames])  # away from the line ends
    expected = 2.0 * a * a * analytic_energy(wl)
    measured = float(np.mean(np.abs(iq) ** 2))
    print(f"[result] E|IQ|^2 measured {measured:.5g} expected {expected:.5g} "
          f"ratio {measured / expected:.4f} over {iq.size} samples", flush=True)
    assert abs(measured / expected - 1.0) < 0.05
    for part in (iq.real, iq.imag):
        assert abs(float(part.mean())) < 0.05 * float(part.std())  # zero mean
        z = (part - part.mean()) / part.std()
        kurt = float(np.mean(z ** 4))
        assert abs(kurt - 3.0) < 0.15, kurt  # Gaussian fourth moment
    # the I and Q variances are equal (circular noise)
    assert abs(float(np.var(iq.real)) / float(np.var(iq.imag)) - 1.0) < 0.05


def test_noise_stream_determinism_and_advance():
    s1, _ = noise_sim(1.0, seed=7)
    s2, _ = noise_sim(1.0, seed=7)
    s3, _ = noise_sim(1.0, seed=8)
    a1, a2, a3 = s1.simulate_frame(), s2.simulate_frame(), s3.simulate_frame()
    assert np.array_equal(a1, a2), "same seed, same frame number: same noise"
    assert rel_l2(a1, a3) > 0.5, "another seed: other noise"
    b1 = s1.simulate_frame()
    assert rel_l2(a1, b1) > 0.5, "the next frame draws new noise"
    s1.set_parameter("noise_seed", "7")  # setting the seed restarts the stream
    assert np.array_equal(s1.simulate_frame(), a1)


def test_noise_adds_linearly_to_the_echo():
    clean, _ = noise_sim(0.0, seed=0, scatterers=3000)
    noisy, _ = noise_sim(0.5, seed=99, scatterers=3000)
    only, _ = noise_sim(0.5, seed=99, scatterers=0)
    c, n, o = clean.simulate_frame(), noisy.simulate_frame(), only.simulate_frame()
    err = rel_l2(n - c, o)
    print(f"[result] (echo + noise) - echo vs noise alone: rel L2 {err:.2e}", flush=True)
    assert err < 1e-5


def test_noise_parameter_round_trip():
    s, _ = noise_sim(0.25, seed=3)
    assert abs(float(s.get_parameter("noise_amplitude")) - 0.25) < 1e-7
    assert s.get_parameter("noise_seed") == "3"


@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_frames_advance_the_noise_stream(use_graph):
    """A batched frame draws the noise of the simulator's own next frame: two
    consecutive batched frames (the second a graph replay when graphs are on)
    differ, and equal frames 1 and 2 of an identical simulator run alone; a frame
    of the simulator alone after them continues its stream."""
    from openbcsim_amd.batch import Batch
    alone, _ = noise_sim(0.8, seed=21, scatterers=2000)
    batched, _ = noise_sim(0.8, seed=21, scatterers=2000)
    other, _ = noise_sim(0.3, seed=5, scatterers=1000, lines=32)  # a second job in the batch
    a = [alone.simulate_frame() for _ in range(3)]
    b = Batch([batched, other], use_graph=use_graph)
    b1 = b.simulate()[0]
    b2 = b.simulate()[0]
    e1, e2 = rel_l2(b1, a[0]), rel_l2(b2, a[1])
    print(f"[result] graph={use_graph}: batched frame 1 vs alone 1 {e1:.2e}, 2 vs 2 {e2:.2e}, "
          f"1 vs 2 {rel_l2(b1, b2):.2f}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5
    assert rel_l2(b1, b2) > 0.5, "the second batched frame draws new noise"
    b.close()
    b3 = batched.simulate_frame()  # frame 3 of the stream, alone
    assert rel_l2(b3, a[2]) < 1e-5
