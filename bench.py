"""Benchmark of the MI355X simulation path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload volume_lut]

A step is one frame of the hot path (projection + fused convolution/demodulation)
over the workload's resident phantom; the timed region holds K frames enqueued
back to back on the simulator's stream with the IQ written to device memory,
bracketed by device synchronisation.  Prints one JSON line (progress on stderr).

Workload (default): the 3D volume scan quoted "at 1M scatterers" -- 1M fixed
float32 points in a 6x6x11 cm box, 64x64 = 4096 pyramidal lines over +-30 deg,
a 256x64x64 float32 LUT beam profile, FFT length 16384, decimation 2.  Others:
spline_mmode, fixed_sector_2d, and each name with "_phase" appended (phase delay on).

roofline: the projection kernel's average duration is measured with HIP events
on the simulator's stream (store_kernel_details) over K frames.  `traffic` is its
HBM bytes per launch from two rocprofv3 --pmc runs of this script in a child
process (FETCH_SIZE, then WRITE_SIZE), each scaled by what the same counter reads
for a plain copy kernel of known size; achieved = traffic / duration against the
8 TB/s HBM peak.  Beside it, the units that bind the kernel: its own counts of
profile-weight evaluations (LUT record gathers) and LDS float additions per
launch, as rates against ceilings measured on the device by kernels of the same
access shape (openbcsim_amd/ceilings.py), the busy fraction of the texture
addresser and data units (a third counter run), and the per-line streaming model
of the reference's kernel (16 B per scatterer evaluation) as a ratio to the peak.

cpu_baseline: the reference's own CPU algorithm (CpuAlgorithm, compiled from the
reference sources into oracle/_ref) on a bounded sample of the same workload --
the first lines of the scan against all scatterers -- with all host cores: the
median of 5 runs after a warm-up.  The same leg checks the sampled served
requests against the FP64 oracle (a few lines of each).

Further legs in the same JSON line (each can be skipped): `configs` (the 2D sector,
the spline M-mode and the volume, phase delay off and on, each also with the
weight cutoff off), `batched_64_jobs` (64 uneven jobs per call vs batched),
`small_jobs` (256 small jobs, where launches dominate) and `serving` (Poisson
arrivals against 16 resident phantoms: p50/p99 latency; saturated throughput
against per-call loops).
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVAL = {"fixed": 16, "spline": 52}
COPY_CALIBRATION_BYTES = 256 << 20


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str):
    from openbcsim_amd import workloads as W
    makers = {"volume_lut": W.volume_lut, "spline_mmode": W.spline_mmode, "fixed_sector_2d": W.fixed_sector_2d}
    base = name[:-len("_phase")] if name.endswith("_phase") else name
    if base not in makers:
        raise SystemExit(f"unknown workload {name}")
    # "<workload>_phase": the same workload with sub-sample phase delay on
    return makers[base](phase_delay=True) if base != name else makers[base]()


def projection_key(wl) -> str:
    return "fixed_projection_kernel_ms" if wl.fixed else "spline_projection_kernel_ms"


def run_gpu(wl, steps: int, warmup: int, tile=None, bins=None):
    from openbcsim_amd.hip_rt import DeviceBuffer, hip
    from openbcsim_amd.pyrfsim import RfSimulator

    sim = RfSimulator("gpu")
    wl.apply(sim)
    if bins:
        sim.set_parameter("bin_accumulation", bins)
    if tile:
        for key, val in zip(("tile_windows", "tile_lines", "tile_block"), tile.split(",")):
            sim.set_parameter(key, val)
    L, S = sim.output_shape()
    out = DeviceBuffer(L * S * 8)
    for _ in range(warmup):
        sim.simulate_lines_device(out.address, L * S)
    sim.synchronize()
    hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.simulate_lines_device(out.address, L * S)
    sim.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps
    # kernel durations with HIP events on the simulator's stream, same frames again
    sim.set_parameter("store_kernel_details", "on")
    kern, counts = {}, {}
    for _ in range(steps):
        sim.simulate_lines_device(out.address, L * S)
        for key in ("fixed_projection_kernel_ms", "spline_projection_kernel_ms", "kernel_convolve_demodulate_ms",
                    "kernel_memset_ms", "kernel_render_ms"):
            try:
                kern.setdefault(key, []).extend(sim.get_debug_data(key))
            except RuntimeError:
                pass
    sim.set_parameter("store_kernel_details", "off")
    # the projection launch's own counts (the same every frame), from one more frame
    sim.set_parameter("projection_counters", "on")
    sim.simulate_lines_device(out.address, L * S)
    for key in ("projection_weight_evals", "projection_lds_adds", "projection_chunk_line_pairs"):
        counts[key] = float(sum(sim.get_debug_data(key)))
    sim.set_parameter("projection_counters", "off")
    counts["bins"] = "fixed" if "bins=fixed" in sim.get_parameter("launch_plan") else "float"
    # PCIe-inclusive rate (host output), reported beside the device-resident value
    sim.simulate_frame()
    t2 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        sim.simulate_frame()
    t3 = time.perf_counter()
    host_ms = (t3 - t2) * 1e3 / max(1, steps // 2)
    sim.close()
    out.close()
    return L, S, ms_per_step, {k: float(np.mean(v)) for k, v in kern.items() if v}, host_ms, counts


def run_configs(steps: int = 5):
    """Every single-simulator config with phase delay off and on (device-resident IQ,
    frame time over `steps` frames after at least 0.1 s of warm-up frames): the 2D sector (100k points,
    128 lines), the spline M-mode (200k cubic splines, 256 lines, one time per
    line) and the 1M-point volume, with the library's choice of LDS bins; each
    again with fixed-point and with float LDS bins forced, and with weight_cutoff 0
    (every pair whose weight is not exactly zero, as the reference sums them)."""
    from openbcsim_amd import workloads as W
    from openbcsim_amd.hip_rt import DeviceBuffer
    from openbcsim_amd.pyrfsim import RfSimulator
    out = {}
    for name, make in (("fixed_sector_2d", W.fixed_sector_2d), ("spline_mmode", W.spline_mmode),
                       ("volume_lut", W.volume_lut)):
        for phase in (False, True):
            wl = make(phase_delay=phase)
            sim = RfSimulator("gpu")
            wl.apply(sim)
            L, S = sim.output_shape()
            buf = DeviceBuffer(L * S * 8)
            res = {}
            for variant, params in (("default", {}), ("fixed_bins", {"bin_accumulation": "fixed"}),
                                    ("float_bins", {"bin_accumulation": "float"}),
                                    ("cutoff_off", {"bin_accumulation": "auto", "weight_cutoff": "0"})):
                for key, val in params.items():
                    sim.set_parameter(key, val)
                # warm-up: at least 3 frames (the block order of a launch shape settles after
                # its first) and 0.1 s of frames (after the idle host-side set-up of a new
                # simulator the first ~10 frames run ~2 % slower)
                n_warm, t_warm = 0, time.perf_counter() + 0.1
                while n_warm < 3 or time.perf_counter() < t_warm:
                    sim.simulate_lines_device(buf.address, L * S)
                    n_warm += 1
                    if n_warm % 4 == 0:
                        sim.synchronize()
                sim.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    sim.simulate_lines_device(buf.address, L * S)
                sim.synchronize()
                res[variant] = (time.perf_counter() - t0) / steps
            dt = res["default"]
            out[f"{name}{'_phase' if phase else ''}"] = {
                "lines": L, "scatterers": wl.num_scatterers, "ms_per_frame": round(dt * 1e3, 4),
                "iq_lines_per_s": round(L / dt, 1), "scatterer_evals_per_s": round(L * wl.num_scatterers / dt, 1),
                "ms_per_frame_fixed_bins": round(res["fixed_bins"] * 1e3, 4),
                "ms_per_frame_float_bins": round(res["float_bins"] * 1e3, 4),
                "ms_per_frame_cutoff_off": round(res["cutoff_off"] * 1e3, 4),
                "iq_lines_per_s_cutoff_off": round(L / res["cutoff_off"], 1)}
            log(f"config {name}{'_phase' if phase else ''}: {dt * 1e3:.4f} ms per frame "
                f"(fixed bins {res['fixed_bins'] * 1e3:.4f}, float bins {res['float_bins'] * 1e3:.4f}, "
                f"{res['cutoff_off'] * 1e3:.4f} without the weight cutoff)")
            sim.close()
            buf.close()
    return out


def small_jobs(num_jobs: int = 256, seed: int = 9):
    """Many small simulations, where per-call launch and synchronisation costs
    dominate: 32-64 sector lines, 10k-50k fixed points or cubic splines, lines of
    2.8 or 6 cm (FFT 2048 or 4096), Gaussian profile."""
    from openbcsim_amd import workloads as W
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(num_jobs):
        n = int(rng.integers(10_000, 50_001))
        lines = int(rng.integers(32, 65))
        length = float(rng.choice([0.028, 0.06]))
        sub_seed = int(rng.integers(0, 1 << 30))
        if rng.integers(0, 2):
            jobs.append(W.spline_mmode(n, lines, seed=sub_seed, line_length=length, decimation=2))
        else:
            jobs.append(W.fixed_sector_2d(n, lines, seed=sub_seed, line_length=length, decimation=2))
    return jobs


def run_batched(num_jobs: int = 64, scale: float = 1.0, reps: int = 3, jobs=None, direct: bool = True):
    """The 64 uneven jobs (scatterer counts log-uniform in [10k, 2M], fixed and spline,
    32-512 lines, FFT 2048-16384, analytic or LUT), or `jobs`: one frame of every job
      - per_call_direct: the plain per-call GPU path -- one synchronous frame per job
        (the reference's call pattern) with the reference's projection method, one
        thread per (scatterer, line) pair accumulating with global atomics
        (projection_mode "direct"), and this library's fused convolution;
      - per_call: one synchronous frame per job with the culled tile kernel and the
        heuristic launch shapes (this library's own per-call path: the batching figure);
      - batched: one batched frame (tuned shapes, persistent launches, graph replay).
    Every leg writes device-resident IQ."""
    from openbcsim_amd import workloads as W
    from openbcsim_amd.batch import Batch
    from openbcsim_amd.hip_rt import DeviceBuffer, hip
    from openbcsim_amd.pyrfsim import RfSimulator

    if jobs is None:
        jobs = W.batched_jobs(num_jobs, seed=3, scale=scale)
    num_jobs = len(jobs)
    sims = []
    for wl in jobs:
        s = RfSimulator("gpu")
        wl.apply(s)
        sims.append(s)
    t0 = time.perf_counter()
    b = Batch(sims)
    total, layout = b.layout()  # plans the batch and tunes every job's shape class
    tune_s = time.perf_counter() - t0
    out = DeviceBuffer(total * 8)
    lines = sum(l for (_, l, _) in layout)
    scat = sum(wl.num_scatterers * wl.num_lines for wl in jobs)

    def per_call():
        for s, (o, l, n) in zip(sims, layout):
            s.simulate_lines_device(out.address + 8 * o, l * n)
            s.synchronize()

    def batched():
        b.simulate_device(out.address, total)
        b.synchronize()

    def set_mode(mode):
        for s in sims:
            s.set_parameter("projection_mode", mode)

    res = {}
    legs = (("per_call_direct", per_call),) if direct else ()
    for name, fn in legs + (("per_call", per_call), ("batched", batched)):
        set_mode("direct" if name == "per_call_direct" else "tiled")
        fn()
        fn()
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        dt = (time.perf_counter() - t0) / reps
        res[name] = dt
    out.close()
    b.close()
    for s in sims:
        s.close()
    out = {"jobs": num_jobs, "lines_per_frame": lines, "scatterer_evals_per_frame": scat,
           "per_call_ms": round(res["per_call"] * 1e3, 3), "batched_ms": round(res["batched"] * 1e3, 3),
           "per_call_lines_per_s": round(lines / res["per_call"], 1),
           "batched_lines_per_s": round(lines / res["batched"], 1),
           "speedup_vs_tiled_per_call": round(res["per_call"] / res["batched"], 3),
           "plan_and_tune_s": round(tune_s, 3)}
    if direct:
        out.update({"per_call_direct_ms": round(res["per_call_direct"] * 1e3, 3),
                    "per_call_direct_lines_per_s": round(lines / res["per_call_direct"], 1),
                    "speedup_vs_plain_per_call": round(res["per_call_direct"] / res["batched"], 3)})
    log(f"batch of {num_jobs} jobs: per-call {res['per_call'] * 1e3:.3f} ms, batched {res['batched'] * 1e3:.3f} ms")
    return out


def run_serving(seconds: float = 60.0, rate: float = 200.0, scale: float = 1.0, seed: int = 11,
                compare_seconds: float = 10.0):
    """Serving loop: 16 resident phantoms (spline M-mode, fixed 2D, 1M-point LUT
    volumes, uneven batched-mix jobs), geometries of 32-512 sector lines and the
    4096-line pyramid; Poisson arrivals at ``rate`` requests/s, each naming a
    phantom, a geometry and a time t, run twice: for ``seconds`` into pinned
    output buffers of the server's (the reported p50/p99) and for
    ``compare_seconds`` into fresh numpy arrays.  Every (phantom, geometry) pair is requested once before the timed
    runs (shape tuning and buffer set-up happen there).  About 1% of the timed
    requests are re-simulated alone on a single simulator and compared (rel L2)."""
    from openbcsim_amd import workloads as W
    from openbcsim_amd.pyrfsim import RfSimulator
    from openbcsim_amd.serve import Server

    phantoms, geometries, pairs = W.serving_mix(seed=5, scale=scale)
    sims = []
    for wl in phantoms:
        s = RfSimulator("gpu")
        wl.apply(s)
        sims.append(s)
    srv = Server()
    for s in sims:
        srv.add_phantom(s)
    for (o, d, l, length) in geometries:
        srv.add_geometry(o, d, l, length)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    warm = [srv.submit(p, g, 0.5) for p, g in pairs]
    for h in warm:
        srv.wait(h)
    warm_s = time.perf_counter() - t0
    log(f"serving: {len(pairs)} phantom/geometry pairs warm; {seconds:.0f} s of Poisson arrivals at {rate:.0f}/s")
    pinned = poisson_run(srv, phantoms, geometries, pairs, rng, rate, seconds, pinned_outputs=True)
    log(f"serving: p50 {pinned['p50_latency_ms']} ms, p99 {pinned['p99_latency_ms']} ms; numpy outputs next")
    pageable = poisson_run(srv, phantoms, geometries, pairs, rng, rate, min(seconds, compare_seconds),
                           pinned_outputs=False)
    log("serving: saturated bursts")
    sat = saturated_serving(srv, sims, geometries, pairs, rng)
    srv.close()
    for s in sims:
        s.close()
    out = dict(pinned)
    samples = out.pop("_samples")
    pageable.pop("_samples")
    out.update({"outputs": "pinned (Server.alloc_output: IQ written by DMA, no copy in wait)",
                "pageable_outputs": {k: pageable[k] for k in ("requests", "p50_latency_ms", "p99_latency_ms",
                                                              "mean_latency_ms", "latency_by_lines",
                                                              "copy_to_caller_ms", "sampled_rel_err_vs_single")},
                "phantoms": len(phantoms), "geometries": len(geometries), "pairs": len(pairs),
                "warmup_s": round(warm_s, 2), "saturated": sat})
    return out, (phantoms, geometries, samples)


def poisson_run(srv, phantoms, geometries, pairs, rng, rate, seconds, pinned_outputs):
    """One timed run of Poisson arrivals; latency = submit to the IQ in the
    caller's buffer.  pinned_outputs: the caller's buffers come from
    Server.alloc_output (recycled after each wait), otherwise fresh numpy arrays."""
    import threading
    from openbcsim_amd import workloads as W
    from openbcsim_amd.pyrfsim import RfSimulator
    srv.set_parameter("reset_stats", 1)
    n = max(1, int(rng.poisson(rate * seconds)))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    choice = rng.integers(len(pairs), size=n)
    times = rng.uniform(0.0, 1.0, n).astype(np.float32)
    sample = set(int(i) for i in rng.choice(n, max(1, n // 100), replace=False))
    kept = {}
    pool, pool_lock, allocated = {}, threading.Lock(), []

    def take(p, g):
        if not pinned_outputs:
            return None
        with pool_lock:
            free = pool.setdefault((p, g), [])
            if free:
                return free.pop()
        arr = srv.alloc_output(p, g)
        arr.fill(0)  # touch the pages before the timed use
        allocated.append(arr)
        return arr

    if pinned_outputs:  # two buffers per pair beforehand, more on demand
        for p, g in pairs:
            bufs = [take(p, g), take(p, g)]
            pool.setdefault((p, g), []).extend(bufs)
    handles = [None] * n
    lat_by = {}
    ready = threading.Semaphore(0)

    def waiter():
        for i in range(n):
            ready.acquire()
            h = handles[i]
            srv.wait(h)
            p, g = pairs[int(choice[i])]
            lat_by.setdefault(int(h.out.shape[0]), []).append((h.latency_ms, h.out.nbytes))
            if i in sample:
                kept[i] = h.out.copy()
            if pinned_outputs:
                with pool_lock:
                    pool[(p, g)].append(h.out)

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    late = 0.0
    for i in range(n):
        now = time.perf_counter() - start
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        else:
            late = max(late, now - arrivals[i])
        p, g = pairs[int(choice[i])]
        handles[i] = srv.submit(p, g, float(times[i]), out=take(p, g))
        ready.release()
    th.join()
    wall = time.perf_counter() - start
    lat = srv.stats("latency_ms")
    copy_ms = srv.stats("copy_ms")
    jobs = srv.stats("frame_jobs")
    lines = float(srv.stats("lines_completed")[0])
    latency_by_lines = {str(k): {"requests": len(v),
                                 "p50_ms": round(float(np.percentile([x[0] for x in v], 50)), 3),
                                 "p99_ms": round(float(np.percentile([x[0] for x in v], 99)), 3),
                                 "iq_mb": round(float(np.mean([x[1] for x in v])) / 1e6, 2)}
                        for k, v in sorted(lat_by.items())}
    # ~1% of the requests again, alone on a single simulator
    worst = 0.0
    for i in sorted(sample):
        p, g = pairs[int(choice[i])]
        s = RfSimulator("gpu")
        W.at_time(phantoms[p], geometries[g], float(times[i])).apply(s)
        ref = s.simulate_frame()
        s.close()
        worst = max(worst, float(np.linalg.norm(kept[i] - ref) / max(np.linalg.norm(ref), 1e-30)))
    for arr in allocated:
        srv.free_output(arr)
    samples = [(pairs[int(choice[i])], float(times[i]), kept[i]) for i in sorted(sample)]
    return {"_samples": samples, "requests": n, "seconds": round(wall, 3), "offered_rate_per_s": rate,
            "achieved_rate_per_s": round(n / wall, 1), "max_submit_lag_ms": round(late * 1e3, 2),
            "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_latency_ms": round(float(np.mean(lat)), 3),
            "latency_by_lines": latency_by_lines,
            "copy_to_caller_ms": {"p50": round(float(np.percentile(copy_ms, 50)), 3),
                                  "p99": round(float(np.percentile(copy_ms, 99)), 3)},
            "iq_lines_per_s": round(lines / wall, 1), "frames": int(len(jobs)),
            "mean_requests_per_frame": round(float(np.mean(jobs)), 2),
            "sampled_rel_err_vs_single": {"requests": int(len(sample)), "max": worst}}


def saturated_serving(srv, sims, geometries, pairs, rng, n: int = 300, clients: int = 8):
    """Throughput with every request queued at once: the server (continuous
    batching; `clients` threads wait for the requests, as concurrent clients do)
    against per-call loops over the same requests -- one simulator per phantom,
    per request set_scan_sequence (the geometry, every timestamp t) and a
    synchronous simulate_lines into host memory, the reference's call pattern
    (MainWindow.cpp:666-695) -- with the reference's projection method (one
    thread per scatterer-line pair, "direct") and with the culled tile kernel.
    The served time is the median of three bursts after a first one.
    Measured twice, each time with the same kind of output buffers for the
    server and the per-call loops, allocated and touched beforehand: pinned
    buffers of the server's (Server.alloc_output; the reported figures) and
    numpy arrays."""
    import threading
    reqs = [(*pairs[int(rng.integers(len(pairs)))], float(rng.uniform(0, 1))) for _ in range(n)]
    lines = sum(geometries[g][0].shape[0] for _, g, _ in reqs)

    def measure(pinned):
        def buffer(shape):
            buf = srv.alloc_output(shape=shape) if pinned else np.empty(shape, np.complex64)
            buf.fill(0)
            return buf

        outs = [buffer(srv.output_shape(p, g)) for p, g, _ in reqs]

        def burst():
            srv.set_parameter("reset_stats", 1)
            t0 = time.perf_counter()
            handles = [srv.submit(p, g, t, out=o) for (p, g, t), o in zip(reqs, outs)]

            def client(k):
                for h in handles[k::clients]:
                    srv.wait(h)

            threads = [threading.Thread(target=client, args=(k,)) for k in range(clients)]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
            return time.perf_counter() - t0

        cold = burst()    # the first burst of this depth also sizes the server's pinned staging pool
        served = float(np.median([burst() for _ in range(3)]))
        jobs = srv.stats("frame_jobs")
        shapes = {}

        def loop(mode):
            for s in sims:
                s.set_parameter("projection_mode", mode)
            t0 = time.perf_counter()
            for p, g, t in reqs:
                o, d, l, length = geometries[g]
                sims[p].set_scan_sequence(o, d, length, l, np.full(o.shape[0], t, np.float32), True)
                key = (p, g)
                if key not in shapes:
                    shapes[key] = buffer(sims[p].output_shape())
                sims[p].simulate_frame(out=shapes[key])
            return time.perf_counter() - t0

        loop("tiled")  # set-up of every (phantom, geometry) shape and output buffer
        tiled = loop("tiled")
        direct = loop("direct")
        for s in sims:
            s.set_parameter("projection_mode", "tiled")
        if pinned:
            for buf in outs + list(shapes.values()):
                srv.free_output(buf)
        return {"served_lines_per_s": round(lines / served, 1), "first_burst_lines_per_s": round(lines / cold, 1),
                "frames": int(len(jobs)), "mean_requests_per_frame": round(float(np.mean(jobs)), 2),
                "per_call_lines_per_s": round(lines / tiled, 1),
                "per_call_direct_lines_per_s": round(lines / direct, 1),
                "speedup_vs_plain_per_call": round(direct / served, 3),
                "speedup_vs_tiled_per_call": round(tiled / served, 3)}

    res = measure(pinned=True)
    res.update({"requests": n, "lines": lines, "clients": clients, "outputs": "pinned",
                "numpy_outputs": measure(pinned=False)})
    return res


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def omp_max_threads() -> int:
    """omp_get_max_threads() of the OpenMP runtime the CPU algorithm runs on."""
    for name in ("libgomp.so.1", "libomp.so"):
        try:
            return int(C.CDLL(name).omp_get_max_threads())
        except (OSError, AttributeError):
            continue
    return 0


def cpu_baseline(wl, target_s: float = 20.0, reps: int = 5):
    """Reference CPU algorithm on the first lines of the scan (all scatterers): one
    warm-up run, then the median of `reps` runs, each about target_s / (reps + 1)
    seconds of wall clock; configuration is outside the timed region."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from backends import Reference, ref_available, Oracle
    import copy

    if ref_available():
        kind, make = "reference", Reference
    else:
        kind, make = "port", Oracle

    def run(n_lines):
        sub = copy.copy(wl)
        o, d, l, t, length = wl.scan
        sub.scan = (o[:n_lines], d[:n_lines], l[:n_lines], t[:n_lines], length)
        sim = make()
        sub.apply(sim)
        t0 = time.perf_counter()
        sim.simulate_frame()
        dt = time.perf_counter() - t0
        sim.close()
        return dt

    nproc = os.cpu_count() or 1
    threads = omp_max_threads() or nproc
    probe = max(1, min(threads, wl.num_lines))
    per_line = run(probe) / probe
    n = int(max(probe, min(wl.num_lines, (target_s / (reps + 1)) / max(per_line, 1e-9))))
    n = max(threads, (n // threads) * threads) if n >= threads else n
    n = min(n, wl.num_lines)
    log(f"cpu baseline ({kind}): {n} lines per run, {threads} threads")
    run(n)  # warm-up
    times = [run(n) for _ in range(reps)]
    t = float(np.median(times))
    N = wl.num_scatterers
    return {"value": n / t, "unit": "IQ lines/s", "cores": threads, "kind": kind,
            "sample": f"first {n} of {wl.num_lines} lines x all {N} scatterers; median of {reps} runs "
                      f"({t:.2f} s each, min {min(times):.2f}, max {max(times):.2f}) after one warm-up run; "
                      f"OpenMP over {threads} threads",
            "nproc": nproc, "omp_max_threads": threads, "cpu_model": cpu_model(),
            "scatterer_evals_per_s": n * N / t, "ns_per_scatterer_per_line": t / (n * N) * 1e9}


def served_vs_oracle(phantoms, geometries, samples, lines_per_request: int = 8):
    """The sampled served requests against the FP64 oracle: for each, the frame of
    its phantom with its scan at its time t (every pair, no weight cutoff), on
    `lines_per_request` evenly spaced lines of the request."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from backends import Oracle, rel_l2
    from openbcsim_amd import workloads as W
    errs = []
    for (p, g), t, out in samples:
        wl = W.at_time(phantoms[p], geometries[g], t)
        o, d, l, ts, length = wl.scan
        sel = np.unique(np.linspace(0, o.shape[0] - 1, min(lines_per_request, o.shape[0])).astype(int))
        wl.scan = (o[sel], d[sel], l[sel], ts[sel], length)
        orc = Oracle()
        wl.apply(orc)
        errs.append(rel_l2(out[sel], orc.simulate_frame()))
        orc.close()
    return {"requests": len(errs), "lines_per_request": lines_per_request,
            "max": float(max(errs)) if errs else None, "median": float(np.median(errs)) if errs else None}


def pmc_child(workload: str, frames: int = 3) -> None:
    """What the counter runs profile: a plain copy kernel of known size (the
    calibration), then `frames` frames of the workload."""
    from openbcsim_amd import ceilings
    from openbcsim_amd.hip_rt import DeviceBuffer
    from openbcsim_amd.pyrfsim import RfSimulator
    log("counter child: copy calibration")
    ceilings.copy_rate(COPY_CALIBRATION_BYTES)  # two dispatches of copy_kernel
    wl = build_workload(workload)
    sim = RfSimulator("gpu")
    wl.apply(sim)
    L, S = sim.output_shape()
    out = DeviceBuffer(L * S * 8)
    for _ in range(frames):
        sim.simulate_lines_device(out.address, L * S)
    sim.synchronize()
    sim.close()
    out.close()
    log(f"counter child: {frames} frames of {workload} done")


def _pmc_values(out_dir: str):
    """{counter: {kernel name: [value per dispatch, in dispatch order]}} from every
    rocprofv3 CSV under out_dir that holds counter values."""
    per = {}
    for path in glob.glob(os.path.join(out_dir, "**", "*.csv"), recursive=True):
        with open(path, newline="") as f:
            rows = [{k.lower().replace("-", "_"): v for k, v in r.items() if k} for r in csv.DictReader(f)]
        if not rows or "counter_value" not in rows[0] or "kernel_name" not in rows[0]:
            continue
        rows.sort(key=lambda r: int(r.get("dispatch_id") or 0))
        for r in rows:
            per.setdefault(r.get("counter_name", "?"), {}).setdefault(r["kernel_name"], []).append(
                float(r["counter_value"]))
    return per


# counter passes of measure_traffic: FETCH_SIZE takes 3 TCC counters, WRITE_SIZE 2, so
# they run apart; the SQ instruction counts ride along with the first
PMC_PASSES = (("FETCH_SIZE", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD"), ("WRITE_SIZE",),
              ("TD_TD_BUSY_sum", "TA_TA_BUSY_sum", "GRBM_GUI_ACTIVE"))
# MI355X: GRBM_GUI_ACTIVE adds up the active cycles of its 8 accelerator complex dies
XCDS = 8


def measure_traffic(workload: str, keep_dir=None):
    """HBM bytes per launch of the projection and convolution kernels, from two
    rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE) on
    `workload`, each scaled by the counter's reading for copy_kernel moving
    COPY_CALIBRATION_BYTES; with the projection's VALU, LDS and vector-memory
    instruction counts, and from a third run the busy cycles of the texture
    addresser and data units beside the active cycles.  Returns (dict, None) or
    (None, reason)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    readings = {}
    unit_err = None
    for npass, counters in enumerate(PMC_PASSES):
        if keep_dir:
            os.makedirs(keep_dir, exist_ok=True)
        out_dir = os.path.abspath(tempfile.mkdtemp(prefix="pmc_", dir=keep_dir or tempfile.gettempdir()))
        cmd = [exe, "--pmc", *counters, "-f", "csv", "-d", out_dir, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--pmc-child", "--workload", workload]
        log(f"counter run: {' '.join(counters)}")
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240,
                               cwd=tempfile.gettempdir(), env=dict(os.environ, TMPDIR=tempfile.gettempdir()))
        except subprocess.TimeoutExpired:
            if npass >= 2:  # the unit-busy pass is optional
                unit_err = f"rocprofv3 --pmc {counters[0]} timed out"
                continue
            return None, f"rocprofv3 --pmc {counters[0]} timed out"
        with open(os.path.join(out_dir, "run.log"), "w") as f:
            f.write(" ".join(cmd) + "\n" + r.stdout)
        vals = _pmc_values(out_dir)
        if counters[0] not in vals:
            files = sorted(os.path.relpath(p, out_dir) for p in glob.glob(os.path.join(out_dir, "**"), recursive=True))
            err = (f"rocprofv3 --pmc {counters[0]}: no counter values (rc {r.returncode}; files {files[:12]}; "
                   f"output head {r.stdout[:600]!r} ... tail {r.stdout[-600:]!r})")
            if npass >= 2:
                unit_err = err
                continue
            return None, err
        readings.update(vals)
        if not keep_dir:
            shutil.rmtree(out_dir, ignore_errors=True)

    def pick(counter, substr):
        got = [v for k, vs in readings.get(counter, {}).items() if substr in k for v in vs]
        return float(np.median(got)) if got else None

    res = {}
    scale = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = pick(counter, "copy_kernel")
        if not cal:
            return None, f"no copy_kernel dispatch in the {counter} run"
        scale[counter] = COPY_CALIBRATION_BYTES / (cal * 1024.0)  # bytes per KiB the counter reports
    for name, substr in (("projection", "project_tile_kernel"), ("convolution", "convolve_demod_kernel")):
        f, w = pick("FETCH_SIZE", substr), pick("WRITE_SIZE", substr)
        if f is None or w is None:
            continue
        res[name] = {"fetch_kib": f, "write_kib": w,
                     "bytes": f * 1024.0 * scale["FETCH_SIZE"] + w * 1024.0 * scale["WRITE_SIZE"]}
        for c in PMC_PASSES[0][1:] + PMC_PASSES[2]:
            v = pick(c, substr)
            if v is not None:
                res[name][c] = v
    res["scale"] = {"fetch": scale["FETCH_SIZE"], "write": scale["WRITE_SIZE"],
                    "calibration": f"copy_kernel moving {COPY_CALIBRATION_BYTES >> 20} MiB read + "
                                   f"{COPY_CALIBRATION_BYTES >> 20} MiB written"}
    if unit_err:
        res["unit_counters_error"] = unit_err
    return res, None


def roofline_object(wl, kind, L, proj_ms, counts, traffic, traffic_err, ceilings, fixed_bins=False):
    """The dominant kernel against HBM (measured bytes) and against the measured
    ceilings of the units its counters show busy: LDS atomics, LUT record
    gathers and spline control-point loads (the launch's own counts), VALU issue (SQ_INSTS_VALU), and the texture
    addresser and data units (their busy cycles per CU over the kernel's active
    cycles, from the counter run)."""
    N = wl.num_scatterers
    t = proj_ms * 1e-3
    rl = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": "project_tile_kernel",
          "kernel_ms": round(proj_ms, 4)}
    proj = (traffic or {}).get("projection")
    if proj:
        b = proj["bytes"]
        rl.update({"achieved": round(b / t / 1e9, 1), "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4),
                   "traffic": round(b),
                   "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this run (child process), "
                                     "scaled by a copy kernel of known size",
                   "traffic_detail": traffic})
    else:
        rl.update({"achieved": None, "frac": None, "traffic": None, "traffic_error": traffic_err})
    model = BYTES_PER_EVAL[kind] * L * N
    rl["streaming_model_bytes"] = model
    rl["streaming_model_ratio"] = round(model / t / 1e9 / HBM_PEAK_GBS, 3)
    units = {}
    if ceilings:
        adds, evals = (counts or {}).get("projection_lds_adds"), (counts or {}).get("projection_weight_evals")
        if adds:
            # the launch's bins: float (ds_add_f32) or 64-bit fixed point (ds_add_u64)
            name, key = (("lds_int64_atomics", "lds_int64_atomics_per_s") if fixed_bins else
                         ("lds_float_atomics", "lds_float_atomics_per_s"))
            c = ceilings[key]
            units[name] = {"per_launch": adds, "achieved_per_s": adds / t, "ceiling_per_s": c,
                           "frac": round(adds / t / c, 4)}
        if evals and wl.profile[0] == "lut":
            c = ceilings["lut_gathers_per_s"]
            units["lut_record_gathers"] = {"per_launch": evals, "achieved_per_s": evals / t, "ceiling_per_s": c,
                                           "frac": round(evals / t / c, 4)}
        pairs = (counts or {}).get("projection_chunk_line_pairs")
        if pairs and kind == "spline" and ceilings.get("spline_wave_hits_per_s"):
            # each listed (chunk, line) pair is one wave hit: its splines' active control points
            c = ceilings["spline_wave_hits_per_s"]
            units["spline_control_point_loads"] = {"per_launch": pairs, "achieved_per_s": pairs / t,
                                                   "ceiling_per_s": c, "frac": round(pairs / t / c, 4)}
        if proj and proj.get("SQ_INSTS_VALU"):
            v, c = proj["SQ_INSTS_VALU"], ceilings["valu_wave_instructions_per_s"]
            units["valu"] = {"per_launch": v, "achieved_per_s": v / t, "ceiling_per_s": c, "frac": round(v / t / c, 4)}
        if proj:
            c = ceilings["copy_bytes_per_s"]
            units["hbm"] = {"per_launch": proj["bytes"], "achieved_per_s": proj["bytes"] / t, "ceiling_per_s": c,
                            "frac": round(proj["bytes"] / t / c, 4)}
        if proj and proj.get("GRBM_GUI_ACTIVE") and ceilings.get("compute_units"):
            # busy cycles of each CU's texture addresser / data unit against the kernel's
            # active cycles per CU (GRBM_GUI_ACTIVE sums the XCDs' clocks)
            active = proj["GRBM_GUI_ACTIVE"] / XCDS
            for name, key in (("texture_addresser", "TA_TA_BUSY_sum"), ("texture_data", "TD_TD_BUSY_sum")):
                if proj.get(key):
                    units[name] = {"busy_cycles": proj[key], "active_cycles_per_cu": active,
                                   "compute_units": ceilings["compute_units"],
                                   "frac": round(proj[key] / (ceilings["compute_units"] * active), 4)}
        rl["units"] = units
        rl["ceilings"] = {k: round(v, 1) for k, v in ceilings.items()}
        if units:
            # the unit nearest its ceiling (not necessarily the one that bounds the
            # kernel: the units overlap, and a unit made faster may not move it)
            rl["busiest_unit"] = max(units, key=lambda k: units[k]["frac"])
    return rl


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="volume_lut")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job and small-job batched-vs-per-call legs")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="scatterer-count scale of the 64 jobs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the per-config (phase off/on) leg")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure the projection's HBM traffic")
    ap.add_argument("--pmc-dir", default=None, help="keep the counter runs' rocprofv3 output under this directory")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the timed serving run (the numpy-output comparison runs 10 s at most)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--tile", default=None, help="projection launch shape 'windows,lines,block' (default: automatic)")
    ap.add_argument("--bins", default=None, choices=("auto", "fixed", "float"),
                    help="LDS bin accumulation of the projection (default: the library's automatic choice)")
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(lines shard across GPUs without communication; see the multi-GPU design notes)",
              file=sys.stderr)
        sys.exit(2)

    # the counter runs first, in child processes, while this process has not yet
    # touched the GPU
    traffic, traffic_err = (None, "skipped (--no-traffic)")
    if not args.no_traffic and not args.tile:
        traffic, traffic_err = measure_traffic(args.workload, args.pmc_dir)
        log(f"traffic: {traffic if traffic else traffic_err}")
    wl = build_workload(args.workload)
    log(f"workload {args.workload}: {wl.num_lines} lines x {wl.num_scatterers} scatterers")
    L, S, ms, kern, host_ms, counts = run_gpu(wl, args.steps, args.warmup, args.tile, args.bins)
    log(f"{ms:.4f} ms per frame; kernels {kern}")
    N = wl.num_scatterers
    evals = L * N
    kind = "fixed" if wl.fixed else "spline"
    proj_ms = kern.get(projection_key(wl))
    from openbcsim_amd import ceilings as CE
    ceil = CE.measure()
    ceil["compute_units"] = CE.cu_count()
    log(f"ceilings: {ceil}")
    lds_bins = counts.pop("bins", None)
    roofline = (roofline_object(wl, kind, L, proj_ms, counts, traffic, traffic_err, ceil,
                                fixed_bins=lds_bins == "fixed") if proj_ms else None)
    line = {
        "metric": "IQ lines/s & scatterer-evals/s at 1M scatterers; serving p50/p99 frame latency",
        "value": round(L / (ms * 1e-3), 1),
        "unit": "IQ lines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom, scan and LUT)",
        "config": {"workload": args.workload, "lines": L, "scatterers": N, "samples_per_line": S,
                   "fft_len": wl.fft_len(), "decimation": int(wl.params.get("radial_decimation", 1)),
                   "profile": wl.profile[0], "phase_delay": wl.params.get("phase_delay", "off"),
                   "weight_cutoff": 1e-8, "lds_bins": lds_bins},
        "scatterer_evals_per_s": round(evals / (ms * 1e-3), 1),
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        "projection_counts": counts,
        "pcie_inclusive_lines_per_s": round(L / (host_ms * 1e-3), 1),
        "roofline": roofline,
    }
    if not args.no_configs:
        line["configs"] = run_configs()
    if not args.no_batch:
        line["batched_64_jobs"] = run_batched(64, args.batch_scale)
        line["small_jobs"] = run_batched(jobs=small_jobs(256), direct=False)
    served = None
    if not args.no_serve:
        line["serving"], served = run_serving(args.serve_seconds, args.serve_rate)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl)
        if served:
            log("served requests against the FP64 oracle")
            line["serving"]["sampled_rel_err_vs_oracle"] = served_vs_oracle(*served)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
