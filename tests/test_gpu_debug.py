"""Debug data of a frame (IAlgorithm::get_debug_data, BaseAlgorithm.cpp:89-94):
with store_kernel_details on, every key the reference GPU algorithm records
(GpuAlgorithm.cpp:193-358) and its GUI prints (MainWindow.cpp:385-396) is
present after simulate_lines, on every projection path."""
import numpy as np
import pytest

from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu

GUI_KEYS = ["stream_numbers", "kernel_memset_ms", "fixed_projection_kernel_ms", "spline_projection_kernel_ms",
            "kernel_forward_fft_ms", "kernel_multiply_fft_ms", "kernel_inverse_fft_ms", "kernel_demodulate_ms",
            "kernel_memcpy_ms"]


@pytest.mark.parametrize("case", ["fixed_one_launch", "fixed_and_spline", "device_output"])
def test_every_reference_debug_key_is_recorded(case):
    from openbcsim_amd.hip_rt import DeviceBuffer
    from openbcsim_amd.pyrfsim import RfSimulator
    wl = W.spline_mmode(3000, 16, seed=3)
    wl.fixed = [W.fixed_sector_2d(3000, 16, seed=4).fixed[0]]
    if case == "fixed_one_launch":
        wl = W.fixed_sector_2d(3000, 16, seed=4)
    s = RfSimulator("gpu")
    wl.apply(s)
    s.set_parameter("store_kernel_details", "on")
    print(f"[case] {case}", flush=True)
    if case == "device_output":
        L, S = s.output_shape()
        buf = DeviceBuffer(L * S * 8)
        s.simulate_lines_device(buf.address, L * S)
        s.synchronize()
    else:
        s.simulate_frame()
    L = wl.num_lines
    got = {}
    for key in GUI_KEYS + ["kernel_convolve_demodulate_ms"]:
        if key == "spline_projection_kernel_ms" and case == "fixed_one_launch":
            continue  # a phantom without splines runs no spline projection (as the reference)
        got[key] = np.asarray(s.get_debug_data(key))
        assert got[key].size >= 1, key
        assert np.all(np.isfinite(got[key])) and np.all(got[key] >= 0), key
    assert got["stream_numbers"].size == L and np.all(got["stream_numbers"] == 0)
    assert got["kernel_forward_fft_ms"][0] == got["kernel_convolve_demodulate_ms"][0] > 0
    for key in ("kernel_multiply_fft_ms", "kernel_inverse_fft_ms", "kernel_demodulate_ms"):
        assert np.all(got[key] == 0.0), key  # fused into the convolution kernel
    assert got["fixed_projection_kernel_ms"][0] > 0
    # 0.0 when one launch stores every bin (no fill kernel), else the fill's time
    assert got["kernel_memset_ms"].size == 1
    if case == "device_output":
        assert got["kernel_memcpy_ms"][0] == 0.0
        buf.close()
    else:
        assert got["kernel_memcpy_ms"][0] > 0
    with pytest.raises(RuntimeError, match="invalid debug data key"):
        s.get_debug_data("no_such_key")
    s.close()
