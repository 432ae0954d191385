"""Host logic of the simulator, batch and serving code under a mock HIP runtime
(CPU only, no GPU): buffer lifetimes and extents of every launch and copy are
checked while streams run asynchronously in random interleavings.  The driver
replays the benchmark's call pattern (tests/host_mock/host_logic_driver.cpp);
the mock aborts on the first error (tests/host_mock/mock_hip.cpp).

The same host code is also built with ThreadSanitizer (data races between the
server's client, scheduler and completion threads) and with AddressSanitizer +
UndefinedBehaviorSanitizer (heap errors, leaks at exit, undefined arithmetic).
These are host-only g++ builds: no device code is compiled."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "host_mock")
CSRC = os.path.join(ROOT, "openbcsim_amd", "csrc")
ROCM_INCLUDE = "/opt/rocm/include"

FLAVOURS = {
    "plain": ["-O1"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
    "asan_ubsan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                   "-fno-sanitize-recover=undefined"],
}
# driver arguments: scatterer scale, serving requests, arrival rate (1/s), batch leg, light
ARGS = {"plain": ["0.02", "200", "2000", "1", "0"], "tsan": ["0.005", "40", "2000", "1", "1"],
        "asan_ubsan": ["0.005", "40", "2000", "1", "1"]}


def build(tmp_path_factory, flavour):
    if shutil.which("g++") is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or the HIP headers are not available")
    exe = str(tmp_path_factory.mktemp("host_mock_" + flavour) / "host_logic_driver")
    cmd = ["g++", "-std=c++17", *FLAVOURS[flavour], "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INCLUDE}",
           f"-I{os.path.join(ROOT, 'include')}", "-pthread", "-Wno-deprecated-declarations", "-Wno-unused-result",
           "-o", exe, os.path.join(HERE, "mock_hip.cpp"), os.path.join(HERE, "host_logic_driver.cpp"),
           os.path.join(CSRC, "sim_host.cpp"), os.path.join(CSRC, "batch_host.cpp"),
           os.path.join(CSRC, "serve_host.cpp")]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    if res.returncode != 0 and flavour != "plain" and "sanitize" in res.stdout and "unrecognized" in res.stdout:
        pytest.skip(f"g++ has no {flavour} runtime")
    assert res.returncode == 0, res.stdout[-4000:]
    return exe


@pytest.mark.parametrize("flavour", sorted(FLAVOURS))
def test_host_logic_lifetimes_and_extents(tmp_path_factory, flavour):
    exe = build(tmp_path_factory, flavour)
    env = dict(os.environ, MOCK_DELAY_US="100",
               TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1",
               ASAN_OPTIONS="halt_on_error=1 detect_leaks=1 exitcode=67",
               UBSAN_OPTIONS="halt_on_error=1 print_stacktrace=1")
    res = subprocess.run([exe, *ARGS[flavour]], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=900, env=env)
    print(res.stdout[-3000:])
    assert res.returncode == 0 and "ALL DONE" in res.stdout, res.stdout[-6000:]
    for marker in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "runtime error:",
                   "ERROR: LeakSanitizer"):
        assert marker not in res.stdout, res.stdout[-6000:]
