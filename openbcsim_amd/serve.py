"""Serving loop over resident phantoms (the bcsim_server_* C-ABI).

    from openbcsim_amd.serve import Server
    srv = Server()
    p = srv.add_phantom(sim)                       # a configured RfSimulator
    g = srv.add_geometry(origins, dirs, lats, 0.12)
    req = srv.submit(p, g, t=0.3)                  # from any thread
    iq = srv.wait(req)                             # complex64 [lines][samples]

Requests are admitted by a device-memory budget and batched continuously: the
scheduler launches whatever is admitted as one batched frame while the
previous frame runs.  A request's IQ equals the frame of a simulator of that
phantom whose scan is the geometry with every timestamp t.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check

_fp = C.POINTER(C.c_float)


@dataclass
class Pending:
    id: int
    out: np.ndarray  # complex64 [lines][samples], filled when the request completes
    latency_ms: float = float("nan")


class Server:
    def __init__(self, memory_budget_bytes: int = 0, max_jobs_per_frame: int = 64, autotune: bool = True):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.bcsim_server_create(C.byref(h)))
        self._h = h
        if memory_budget_bytes:
            self.set_parameter("memory_budget_bytes", int(memory_budget_bytes))
        self.set_parameter("max_jobs_per_frame", int(max_jobs_per_frame))
        self.set_parameter("autotune", "on" if autotune else "off")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bcsim_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parameter(self, key: str, value):
        check(self._lib.bcsim_server_set_parameter(self._h, key.encode(), str(value).encode()))

    def add_phantom(self, sim) -> int:
        idx = C.c_int()
        check(self._lib.bcsim_server_add_phantom(self._h, sim.handle, C.byref(idx)))
        return idx.value

    def add_geometry(self, origins, directions, lateral_dirs, line_length: float) -> int:
        arrs = [np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 3)) for a in
                (origins, directions, lateral_dirs)]
        n = arrs[0].shape[0]
        if any(a.shape[0] != n for a in arrs):
            raise ValueError("origins, directions and lateral_dirs must have the same number of lines")
        idx = C.c_int()
        check(self._lib.bcsim_server_add_geometry(self._h, n, *(a.ctypes.data_as(_fp) for a in arrs),
                                                  float(line_length), C.byref(idx)))
        return idx.value

    def output_shape(self, phantom: int, geometry: int):
        nl, ns = C.c_int(), C.c_int()
        check(self._lib.bcsim_server_get_output_shape(self._h, phantom, geometry, C.byref(nl), C.byref(ns)))
        return nl.value, ns.value

    def alloc_output(self, phantom: int = None, geometry: int = None, shape=None) -> np.ndarray:
        """Pinned complex64 array (lines, samples) for submit(out=...): the frame's IQ
        is written into it by DMA, with no staging copy in wait.  Valid until
        free_output(array) or close()."""
        nl, ns = shape if shape is not None else self.output_shape(phantom, geometry)
        n = int(nl) * int(ns)
        p = C.c_void_p()
        check(self._lib.bcsim_server_alloc_host(self._h, n * 8, C.byref(p)))
        flat = np.ctypeslib.as_array((C.c_float * (2 * n)).from_address(p.value))
        return flat.view(np.complex64).reshape(int(nl), int(ns))

    def free_output(self, arr: np.ndarray):
        """Release an array of alloc_output (an error while an unreported request
        writes into it)."""
        check(self._lib.bcsim_server_free_host(self._h, C.c_void_p(arr.__array_interface__["data"][0])))

    def submit(self, phantom: int, geometry: int, t: float = 0.0, out: np.ndarray = None) -> Pending:
        nl, ns = self.output_shape(phantom, geometry)
        if out is None:
            out = np.empty((nl, ns), np.complex64)
        if out.dtype != np.complex64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous complex64 array")  # its size is checked by the library
        rid = C.c_uint64()
        check(self._lib.bcsim_server_submit(self._h, phantom, geometry, float(t), out.ctypes.data_as(_fp),
                                            out.size, C.byref(rid)))
        return Pending(rid.value, out)

    def wait(self, req: Pending, timeout_ms: int = -1) -> np.ndarray:
        nl, ns, lat = C.c_int(), C.c_int(), C.c_double()
        check(self._lib.bcsim_server_wait(self._h, req.id, int(timeout_ms), C.byref(nl), C.byref(ns),
                                          C.byref(lat)))
        req.latency_ms = lat.value
        return req.out.reshape(-1)[: nl.value * ns.value].reshape(nl.value, ns.value)

    def stats(self, key: str):
        n = C.c_size_t()
        check(self._lib.bcsim_server_get_stats(self._h, key.encode(), None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.float64)
        check(self._lib.bcsim_server_get_stats(self._h, key.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                                               n.value, C.byref(n)))
        return out[: n.value]
