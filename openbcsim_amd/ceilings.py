"""Measured ceilings of the units that bound the scatterer projection (the
benchmark's roofline), from the kernels of csrc/ceilings.hip built into
libbcsim_ceilings.so:

* ``lds_add_rate``: ds_add_f32 per second, chip-wide, with the tile kernel's
  pattern (4 rows of 9740 bins per 1024-thread workgroup, each wave adding at
  bins spread over a 256-bin window); mode 2: ds_add_u64 into 2 rows of 64-bit
  bins (the fixed-point bins);
* ``gather_rate``: 32-byte LUT record gathers per second, each wave's lanes in
  a chunk-sized neighbourhood of an L2-resident table (the projection's LUT
  lookups with nothing else in the way);
* ``spline_rate``: spline hits of one wave per second from the control-point
  loads alone (4 active control points of 64 consecutive splines of a
  200k x 10 dataset, [cp][n] float arrays: 12 coalesced loads per lane);
* ``valu_rate``: VALU wave instructions per second (independent FMA chains);
* ``copy_rate``: bytes moved per second by a plain float4 copy (read + write).

Each is timed with HIP events on the device; this is measurement
infrastructure, not part of the simulation path."""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbcsim_ceilings.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m openbcsim_amd.build`")
        lib = C.CDLL(LIB_PATH)
        lib.ceil_lds_add.restype = C.c_float
        lib.ceil_lds_add.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.ceil_gather.restype = C.c_float
        lib.ceil_gather.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.ceil_spline.restype = C.c_float
        lib.ceil_spline.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p]
        lib.ceil_valu.restype = C.c_float
        lib.ceil_valu.argtypes = [C.c_int, C.c_int, C.c_void_p]
        lib.ceil_copy.restype = C.c_float
        lib.ceil_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong]
        lib.ceil_cu_count.restype = C.c_int
        lib.ceil_cu_count.argtypes = []
        _lib = lib
    return _lib


def _ms(v: float, what: str) -> float:
    if not v > 0:
        raise RuntimeError(f"{what} probe failed")
    return float(v)


def lds_add_rate(workgroups: int = 1024, steps: int = 4096, mode: int = 0) -> float:
    from .hip_rt import DeviceBuffer
    sink = DeviceBuffer(workgroups * 1024 * 4)
    ms = _ms(load().ceil_lds_add(workgroups, steps, mode, C.c_void_p(sink.address)), "lds_add")
    sink.close()
    return workgroups * 1024.0 * steps / (ms * 1e-3)


def gather_rate(slices: int = 32, span=(5, 10, 10), workgroups: int = 1024, steps: int = 2048) -> float:
    """Record gathers per second from a table of `slices` x 64 x 64 records of 32
    bytes (4 MiB at 32 slices: L2-resident), lanes in a `span` neighbourhood."""
    from .hip_rt import DeviceBuffer
    table = DeviceBuffer(slices * 64 * 64 * 32)
    sink = DeviceBuffer(workgroups * 1024 * 4)
    ms = _ms(load().ceil_gather(C.c_void_p(table.address), slices, span[0], span[1], span[2], workgroups, steps,
                                C.c_void_p(sink.address)), "gather")
    table.close()
    sink.close()
    return workgroups * 1024.0 * steps / (ms * 1e-3)


def spline_rate(n: int = 200_000, m: int = 10, workgroups: int = 1024, steps: int = 512) -> float:
    """Wave spline hits per second: per wave and step, the 4 active control points
    of 64 consecutive splines (random chunk and knot interval) of an n x m dataset."""
    from .hip_rt import DeviceBuffer
    arrays = [DeviceBuffer(n * m * 4) for _ in range(3)]
    sink = DeviceBuffer(workgroups * 1024 * 4)
    ms = _ms(load().ceil_spline(*(C.c_void_p(a.address) for a in arrays), n, m, workgroups, steps,
                                C.c_void_p(sink.address)), "spline")
    for a in arrays:
        a.close()
    sink.close()
    return workgroups * 16.0 * steps / (ms * 1e-3)


def valu_rate(workgroups: int = 1024, iters: int = 4096) -> float:
    """VALU wave instructions per second: 8 independent FMA chains per thread."""
    from .hip_rt import DeviceBuffer
    sink = DeviceBuffer(workgroups * 1024 * 4)
    ms = _ms(load().ceil_valu(workgroups, iters, C.c_void_p(sink.address)), "valu")
    sink.close()
    return workgroups * 16.0 * iters * 8 / (ms * 1e-3)


def copy_rate(nbytes: int = 1 << 30) -> float:
    from .hip_rt import DeviceBuffer
    src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    ms = _ms(load().ceil_copy(C.c_void_p(dst.address), C.c_void_p(src.address), nbytes // 16), "copy")
    src.close()
    dst.close()
    return 2.0 * nbytes / (ms * 1e-3)


def cu_count() -> int:
    n = load().ceil_cu_count()
    if n < 1:
        raise RuntimeError("compute unit count query failed")
    return int(n)


def measure() -> dict:
    """Every ceiling, in the units the benchmark reports."""
    return {"lds_float_atomics_per_s": lds_add_rate(),
            "lds_float_atomics_per_s_conflict_free": lds_add_rate(mode=1),
            "lds_int64_atomics_per_s": lds_add_rate(mode=2),
            "lut_gathers_per_s": gather_rate(),
            "lut_gathers_per_s_one_line_per_lane": gather_rate(span=(32, 64, 64)),
            "spline_wave_hits_per_s": spline_rate(),
            "valu_wave_instructions_per_s": valu_rate(),
            "copy_bytes_per_s": copy_rate()}
