"""Host logic of the simulator, batch and serving code under a mock HIP runtime
(CPU only, no GPU): buffer lifetimes and extents of every launch and copy are
checked while streams run asynchronously in random interleavings.  The driver
replays the benchmark's call pattern (tests/host_mock/host_logic_driver.cpp);
the mock aborts on the first error (tests/host_mock/mock_hip.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "host_mock")
CSRC = os.path.join(ROOT, "openbcsim_amd", "csrc")
ROCM_INCLUDE = "/opt/rocm/include"


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or the HIP headers are not available")
    exe = str(tmp_path_factory.mktemp("host_mock") / "host_logic_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INCLUDE}",
           f"-I{os.path.join(ROOT, 'include')}", "-pthread", "-Wno-deprecated-declarations", "-Wno-unused-result",
           "-o", exe, os.path.join(HERE, "mock_hip.cpp"), os.path.join(HERE, "host_logic_driver.cpp"),
           os.path.join(CSRC, "sim_host.cpp"), os.path.join(CSRC, "batch_host.cpp"),
           os.path.join(CSRC, "serve_host.cpp")]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-4000:]
    return exe


def test_host_logic_lifetimes_and_extents(driver):
    env = dict(os.environ, MOCK_DELAY_US="100")
    res = subprocess.run([driver, "0.02", "200", "2000", "1"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=600, env=env)
    print(res.stdout[-3000:])
    assert res.returncode == 0 and "ALL DONE" in res.stdout, res.stdout[-3000:]
