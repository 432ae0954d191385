"""A C++ program using the IAlgorithm facade header (include/bcsim/LibBCSim.hpp),
linked against libbcsim_hip.so, checked against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from backends import Oracle, rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openbcsim_amd", "_build", "facade_check")


def test_cxx_facade_program_matches_oracle(tmp_path):
    assert os.path.exists(BIN), "build the facade check first (openbcsim_amd.build.build_cxx_check)"
    out = tmp_path / "frame.bin"
    print("[case] facade_check", flush=True)
    res = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = out.read_bytes()
    hdr = np.frombuffer(raw[:24], np.int32)
    L, S, NF, NS, M, P = (int(x) for x in hdr)
    f = np.frombuffer(raw[24:], np.float32)
    pos = 0

    def take(n, shape):
        nonlocal pos
        a = f[pos:pos + n].reshape(shape)
        pos += n
        return a

    o = take(3 * L, (L, 3)); d = take(3 * L, (L, 3)); l = take(3 * L, (L, 3)); t = take(L, (L,))
    fixed = take(4 * NF, (NF, 4)); knots = take(M + P + 1, (M + P + 1,))
    cps = take(NS * M * 3, (NS, M, 3)); amps = take(NS, (NS,)); exc = take(55, (55,))
    iq = f[pos:pos + 2 * L * S].view(np.complex64).reshape(L, S)
    orc = Oracle()
    orc.set_parameter("sound_speed", "1540")
    orc.set_parameter("radial_decimation", "2")
    orc.set_parameter("phase_delay", "on")
    orc.set_excitation(exc, 27, 50e6, 2.5e6)
    orc.set_scan_sequence(o, d, 0.05, l, t)
    orc.add_fixed_scatterers(fixed)
    orc.add_spline_scatterers(P, knots, cps, amps)
    orc.set_analytical_beam_profile(1e-3, 1.5e-3)
    err = rel_l2(iq, orc.simulate_frame())
    print(f"[result] facade rel_l2={err:.3e}", flush=True)
    assert err <= 1e-4


def test_c99_caller_matches_oracle(tmp_path):
    """tests/cxx/c_caller.c, a C99 program over include/bcsim_c.h."""
    exe = os.path.join(ROOT, "openbcsim_amd", "_build", "c_caller")
    assert os.path.exists(exe), "build the C caller first (openbcsim_amd.build.build_cxx_check)"
    out = tmp_path / "frame.bin"
    print("[case] c_caller", flush=True)
    res = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = out.read_bytes()
    L, S, N = (int(x) for x in np.frombuffer(raw[:12], np.int32))
    f = np.frombuffer(raw[12:], np.float32)
    sizes = [3 * L, 3 * L, 3 * L, L, 4 * N, 55]
    parts, pos = [], 0
    for n in sizes:
        parts.append(f[pos:pos + n])
        pos += n
    o, d, l, t, xyza, exc = parts
    iq = f[pos:pos + 2 * L * S].view(np.complex64).reshape(L, S)
    orc = Oracle()
    orc.set_parameter("radial_decimation", "2")
    orc.set_excitation(exc, 27, 50e6, 2.5e6)
    orc.set_scan_sequence(o.reshape(L, 3), d.reshape(L, 3), 0.04, l.reshape(L, 3), t)
    orc.add_fixed_scatterers(xyza.reshape(N, 4))
    orc.set_analytical_beam_profile(1e-3, 1e-3)
    err = rel_l2(iq, orc.simulate_frame())
    print(f"[result] c_caller rel_l2={err:.3e}", flush=True)
    assert err <= 1e-4
