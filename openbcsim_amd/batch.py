"""Batched frames over many simulators (the bcsim_batch_* C-ABI).

    from openbcsim_amd.batch import Batch
    b = Batch([sim_a, sim_b, ...])        # configured RfSimulator objects (jobs)
    frames = b.simulate()                 # list of complex64 [line][sample] arrays

One frame of every job runs as one persistent projection launch and one
persistent convolution launch (captured into a HIP graph and replayed while no
job changes).  Each job's result equals what its own simulate_frame() returns,
up to float summation order.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib
from ._lib import check


class Batch:
    def __init__(self, sims: Sequence = (), use_graph: bool = True):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.bcsim_batch_create(C.byref(h)))
        self._h = h
        self._sims = []
        self.set_parameter("use_graph", "on" if use_graph else "off")
        for s in sims:
            self.add(s)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bcsim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parameter(self, key: str, value: str):
        check(self._lib.bcsim_batch_set_parameter(self._h, key.encode(), str(value).encode()))

    def add(self, sim) -> int:
        idx = C.c_int()
        check(self._lib.bcsim_batch_add(self._h, sim.handle, C.byref(idx)))
        self._sims.append(sim)  # keep the handle alive as long as the batch
        return idx.value

    def clear(self):
        check(self._lib.bcsim_batch_clear(self._h))
        self._sims = []

    def layout(self):
        """(total complex elements, [(offset, lines, samples) per job])."""
        n = len(self._sims)
        total = C.c_size_t()
        nj = C.c_int()
        offs = (C.c_size_t * max(n, 1))()
        lines = (C.c_int * max(n, 1))()
        samples = (C.c_int * max(n, 1))()
        check(self._lib.bcsim_batch_get_layout(self._h, C.byref(total), C.byref(nj), offs, lines, samples, n))
        return total.value, [(offs[i], lines[i], samples[i]) for i in range(nj.value)]

    def simulate(self) -> List[np.ndarray]:
        """One frame of every job into host memory: list of complex64 [line][sample]."""
        total, jobs = self.layout()
        flat = np.empty(max(total, 1), np.complex64)
        check(self._lib.bcsim_batch_simulate(self._h, flat.ctypes.data_as(C.POINTER(C.c_float)), flat.size))
        return [flat[o:o + l * s].reshape(l, s) for (o, l, s) in jobs]

    def simulate_device(self, device_ptr: int, capacity: int):
        check(self._lib.bcsim_batch_simulate_device(self._h, C.c_void_p(device_ptr), capacity))

    def synchronize(self):
        check(self._lib.bcsim_batch_synchronize(self._h))

    def get_debug_data(self, key: str):
        n = C.c_size_t()
        check(self._lib.bcsim_batch_get_debug_data(self._h, key.encode(), None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.float64)
        check(self._lib.bcsim_batch_get_debug_data(self._h, key.encode(),
                                                   out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return list(out[: n.value])
