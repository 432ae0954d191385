"""The allocation-free frame entry point, bcsim_simulate_lines_device_noalloc,
called through the C-ABI: its IQ equals bcsim_simulate_lines_device's (the same
kernels; float summation order aside) on two calls in a row, the second with a
smaller frame while the shared scratch still holds the first frame's bins; the
first frame again afterwards; and its errors (tables not built, a frame larger
than the scratch, a null output)."""
import ctypes as C

import numpy as np
import pytest

from backends import rel_l2
from openbcsim_amd import _lib
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(wl):
    from openbcsim_amd.pyrfsim import RfSimulator
    s = RfSimulator("gpu")
    wl.apply(s)
    return s


def frame(sim, fn):
    """One frame through the named C-ABI entry point into fresh device memory,
    copied back after the handle's stream has drained."""
    from openbcsim_amd.hip_rt import DeviceBuffer
    lib = _lib.load()
    L, S = sim.output_shape()
    buf = DeviceBuffer(L * S * 8)
    nl, ns = C.c_int(), C.c_int()
    rc = getattr(lib, fn)(sim.handle, C.c_void_p(buf.address), L * S, C.byref(nl), C.byref(ns))
    assert rc == 0, lib.bcsim_last_error().decode()
    assert (nl.value, ns.value) == (L, S)
    _lib.check(lib.bcsim_synchronize(sim.handle))
    out = np.empty((L, S), np.complex64)
    buf.to_host(out)
    buf.close()
    return out


def mixed_frame_sim():
    """A fixed set and a spline set with a time per line: two launches accumulate
    into the bins, so they are zeroed first."""
    wl = W.spline_mmode(20000, 96, seed=61, decimation=2)
    wl.fixed = [W.fixed_sector_2d(30000, 96, seed=62).fixed[0]]
    return make(wl)


def smaller_phase_sim():
    """Fewer lines and scatterers, complex bins (phase delay), splines rendered at
    one time (their points also in the scratch), plus a fixed set."""
    wl = W.spline_mmode(6000, 40, seed=63, phase_delay=True, decimation=1, all_timestamps_equal=True,
                        line_length=0.06)
    wl.fixed = [W.fixed_sector_2d(8000, 40, seed=64, line_length=0.06).fixed[0]]
    return make(wl)


def test_noalloc_frames_match_the_allocating_entry_point():
    a = mixed_frame_sim()
    b = smaller_phase_sim()
    print("[case] no-alloc: 96-line mixed frame, then a 40-line phase-delay frame", flush=True)
    ref_a = frame(a, "bcsim_simulate_lines_device")  # builds a's tables
    ref_b = frame(b, "bcsim_simulate_lines_device")
    got_a = frame(a, "bcsim_simulate_lines_device_noalloc")
    got_b = frame(b, "bcsim_simulate_lines_device_noalloc")  # smaller, over a's bins in the scratch
    again_a = frame(a, "bcsim_simulate_lines_device_noalloc")
    e = [rel_l2(got_a, ref_a), rel_l2(got_b, ref_b), rel_l2(again_a, ref_a)]
    print(f"[result] no-alloc vs allocating: first {e[0]:.2e}, second (smaller) {e[1]:.2e}, "
          f"first again {e[2]:.2e}", flush=True)
    assert max(e) <= 1e-6
    assert np.abs(ref_a).max() > 0 and np.abs(ref_b).max() > 0
    # a frame through the allocating path still works after the no-alloc ones
    assert rel_l2(frame(b, "bcsim_simulate_lines_device"), ref_b) <= 1e-6
    a.close()
    b.close()


def test_noalloc_errors():
    from openbcsim_amd.hip_rt import DeviceBuffer
    lib = _lib.load()
    nl, ns = C.c_int(), C.c_int()
    buf = DeviceBuffer(1 << 20)
    # tables not built yet: no frame has run since configuration
    s = make(W.fixed_sector_2d(2000, 8, seed=65))
    rc = lib.bcsim_simulate_lines_device_noalloc(s.handle, C.c_void_p(buf.address), 1 << 17, C.byref(nl),
                                                 C.byref(ns))
    assert rc == 2 and b"tables are not built" in lib.bcsim_last_error()
    frame(s, "bcsim_simulate_lines_device")
    s.set_parameter("radial_decimation", "3")  # a configuration change needs a frame again
    rc = lib.bcsim_simulate_lines_device_noalloc(s.handle, C.c_void_p(buf.address), 1 << 17, C.byref(nl),
                                                 C.byref(ns))
    assert rc == 2
    frame(s, "bcsim_simulate_lines_device")
    rc = lib.bcsim_simulate_lines_device_noalloc(s.handle, None, 1 << 17, C.byref(nl), C.byref(ns))
    assert rc == 1 and b"null output buffer" in lib.bcsim_last_error()
    rc = lib.bcsim_simulate_lines_device_noalloc(s.handle, C.c_void_p(buf.address), 4, C.byref(nl), C.byref(ns))
    assert rc == 5
    s.close()
    # bins beyond the fixed scratch: 4096 lines x 9740 samples x 4 bytes
    v = make(W.volume_lut(2000, seed=66))
    L, S = v.output_shape()
    out = DeviceBuffer(L * S * 8)
    frame(v, "bcsim_simulate_lines_device")
    rc = lib.bcsim_simulate_lines_device_noalloc(v.handle, C.c_void_p(out.address), L * S, C.byref(nl), C.byref(ns))
    assert rc == 5 and b"bytes of scratch" in lib.bcsim_last_error()
    out.close()
    v.close()
    buf.close()
