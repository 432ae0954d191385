"""Benchmark of the MI355X simulation path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload volume_lut]

A step is one frame of the hot path (projection + fused convolution/demodulation)
over the workload's resident phantom; the timed region holds K frames enqueued
back to back on the simulator's stream with the IQ written to device memory,
bracketed by device synchronisation.  Prints one JSON line.

Workload (default): the 3D volume scan quoted "at 1M scatterers" -- 1M fixed
float32 points in a 6x6x11 cm box, 64x64 = 4096 pyramidal lines over +-30 deg,
a 256x64x64 float32 LUT beam profile, FFT length 16384, decimation 2.  Others:
spline_mmode, fixed_sector_2d, and each name with "_phase" appended (phase delay on).

roofline: the projection kernel's average duration is measured with HIP events
on the simulator's stream (store_kernel_details) over the same K frames; its
algorithmic bytes are 16 B per scatterer evaluation (x, y, z, amplitude read
once per line in the per-line streaming model) times L*N evaluations per launch.

cpu_baseline: the reference's own CPU algorithm (CpuAlgorithm, compiled from the
reference sources into oracle/_ref) on a bounded sample of the same workload --
the first lines of the scan against all scatterers -- with all host cores.

Further legs in the same JSON line (each can be skipped): `configs` (the 2D sector,
the spline M-mode and the volume, phase delay off and on), `batched_64_jobs` (64
uneven jobs per call vs batched) and `serving` (Poisson arrivals against 16 resident
phantoms: p50/p99 latency; saturated throughput against per-call loops).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVAL = {"fixed": 16, "spline": 52}
# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the scales a 256 MiB device copy reads back as), for the committed kernels;
# profiles/project_tile_volume_lut_1M_4096_pmc.md gives the commands and counts.
MEASURED_TRAFFIC = {"volume_lut": {"bytes": 1.867e6 * 1024 * 2.0 + 1.573e5 * 1024,
                                   "source": "profiles/project_tile_volume_lut_1M_4096_pmc.md"}}


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


def run_gpu(wl, steps: int, warmup: int, tile=None):
    from openbcsim_amd.hip_rt import DeviceBuffer, hip
    from openbcsim_amd.pyrfsim import RfSimulator

    sim = RfSimulator("gpu")
    wl.apply(sim)
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
    kern = {}
    for _ in range(steps):
        sim.simulate_lines_device(out.address, L * S)
        for key in ("fixed_projection_kernel_ms", "spline_projection_kernel_ms", "kernel_convolve_demodulate_ms",
                    "kernel_memset_ms", "kernel_render_ms"):
            try:
                kern.setdefault(key, []).extend(sim.get_debug_data(key))
            except RuntimeError:
                pass
    sim.set_parameter("store_kernel_details", "off")
    # PCIe-inclusive rate (host output), reported beside the device-resident value
    sim.simulate_frame()
    t2 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        sim.simulate_frame()
    t3 = time.perf_counter()
    host_ms = (t3 - t2) * 1e3 / max(1, steps // 2)
    sim.close()
    out.close()
    return L, S, ms_per_step, {k: float(np.mean(v)) for k, v in kern.items() if v}, host_ms


def run_configs(steps: int = 5):
    """Every single-simulator config with phase delay off and on (device-resident IQ,
    frame time over `steps` frames after one warm-up): the 2D sector (100k points,
    128 lines), the spline M-mode (200k cubic splines, 256 lines, one time per
    line) and the 1M-point volume."""
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
            sim.simulate_lines_device(buf.address, L * S)
            sim.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                sim.simulate_lines_device(buf.address, L * S)
            sim.synchronize()
            dt = (time.perf_counter() - t0) / steps
            out[f"{name}{'_phase' if phase else ''}"] = {
                "lines": L, "scatterers": wl.num_scatterers, "ms_per_frame": round(dt * 1e3, 4),
                "iq_lines_per_s": round(L / dt, 1), "scatterer_evals_per_s": round(L * wl.num_scatterers / dt, 1)}
            sim.close()
            buf.close()
    return out


def run_batched(num_jobs: int = 64, scale: float = 1.0, reps: int = 3):
    """The 64 uneven jobs (scatterer counts log-uniform in [10k, 2M], fixed and spline,
    32-512 lines, FFT 2048-16384, analytic or LUT): one frame of every job
      - per_call_direct: the plain per-call GPU path -- one synchronous frame per job
        (the reference's call pattern) with the reference's projection method, one
        thread per (scatterer, line) pair accumulating with global atomics
        (projection_mode "direct"), and this library's fused convolution;
      - per_call: one synchronous frame per job with the culled tile kernel and the
        heuristic launch shapes;
      - batched: one batched frame (tuned shapes, persistent launches, graph replay).
    Every leg writes device-resident IQ."""
    from openbcsim_amd import workloads as W
    from openbcsim_amd.batch import Batch
    from openbcsim_amd.hip_rt import DeviceBuffer, hip
    from openbcsim_amd.pyrfsim import RfSimulator

    jobs = W.batched_jobs(num_jobs, seed=3, scale=scale)
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
    for name, fn in (("per_call_direct", per_call), ("per_call", per_call), ("batched", batched)):
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
    return {"jobs": num_jobs, "lines_per_frame": lines, "scatterer_evals_per_frame": scat,
            "per_call_direct_ms": round(res["per_call_direct"] * 1e3, 3),
            "per_call_ms": round(res["per_call"] * 1e3, 3), "batched_ms": round(res["batched"] * 1e3, 3),
            "per_call_direct_lines_per_s": round(lines / res["per_call_direct"], 1),
            "per_call_lines_per_s": round(lines / res["per_call"], 1),
            "batched_lines_per_s": round(lines / res["batched"], 1),
            "speedup_vs_plain_per_call": round(res["per_call_direct"] / res["batched"], 3),
            "speedup_vs_tiled_per_call": round(res["per_call"] / res["batched"], 3),
            "plan_and_tune_s": round(tune_s, 3)}


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
    pinned = poisson_run(srv, phantoms, geometries, pairs, rng, rate, seconds, pinned_outputs=True)
    pageable = poisson_run(srv, phantoms, geometries, pairs, rng, rate, min(seconds, compare_seconds),
                           pinned_outputs=False)
    sat = saturated_serving(srv, sims, geometries, pairs, rng)
    srv.close()
    for s in sims:
        s.close()
    out = dict(pinned)
    out.update({"outputs": "pinned (Server.alloc_output: IQ written by DMA, no copy in wait)",
                "pageable_outputs": {k: pageable[k] for k in ("requests", "p50_latency_ms", "p99_latency_ms",
                                                              "mean_latency_ms", "latency_by_lines",
                                                              "copy_to_caller_ms", "sampled_rel_err_vs_single")},
                "phantoms": len(phantoms), "geometries": len(geometries), "pairs": len(pairs),
                "warmup_s": round(warm_s, 2), "saturated": sat})
    return out


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
    return {"requests": n, "seconds": round(wall, 3), "offered_rate_per_s": rate,
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


def cpu_baseline(wl, target_s: float = 12.0):
    """Reference CPU algorithm on the first lines of the scan (all scatterers)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from backends import Reference, ref_available, Oracle
    import copy

    cores = os.cpu_count() or 1
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
        return time.perf_counter() - t0

    probe = max(1, min(cores, wl.num_lines))
    t_probe = run(probe)
    per_line = t_probe / probe
    n = int(max(probe, min(wl.num_lines, (target_s / 3) / max(per_line, 1e-9))))
    n = max(cores, (n // cores) * cores) if n >= cores else n
    n = min(n, wl.num_lines)
    times = [run(n) for _ in range(3)]
    t = float(np.median(times))
    return {"value": n / t, "unit": "IQ lines/s", "cores": cores, "kind": kind,
            "sample": f"first {n} of {wl.num_lines} lines x all {wl.num_scatterers} scatterers, "
                      f"median of 3 runs ({t:.2f} s each), OpenMP over {cores} host threads",
            "scatterer_evals_per_s": n * wl.num_scatterers / t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="volume_lut")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job batched-vs-per-call leg")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="scatterer-count scale of the 64 jobs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the per-config (phase off/on) leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the timed serving run (the numpy-output comparison runs 10 s at most)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--tile", default=None, help="projection launch shape 'windows,lines,block' (default: automatic)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a plain copy kernel of "
                         "known size, to scale rocprof FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration_mib > 0:
        from openbcsim_amd.hip_rt import DeviceBuffer, hip
        nbytes = args.copy_calibration_mib << 20
        src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        hip().hipMemcpy(C.c_void_p(dst.address), C.c_void_p(src.address), nbytes, 3)
        hip().hipDeviceSynchronize()
        src.close()
        dst.close()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(lines shard across GPUs without communication; see the multi-GPU design notes)",
              file=sys.stderr)
        sys.exit(2)

    wl = build_workload(args.workload)
    L, S, ms, kern, host_ms = run_gpu(wl, args.steps, args.warmup, args.tile)
    N = wl.num_scatterers
    evals = L * N
    kind = "fixed" if wl.fixed else "spline"
    pkey = projection_key(wl)
    proj_ms = kern.get(pkey)
    roofline = None
    if proj_ms:
        algo_bytes = BYTES_PER_EVAL[kind] * evals
        achieved = algo_bytes / (proj_ms * 1e-3) / 1e9
        measured = MEASURED_TRAFFIC.get(args.workload) if not args.tile else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": round(measured["bytes"]) if measured else None,
                    "kernel": pkey.replace("_ms", ""), "kernel_ms": round(proj_ms, 4),
                    "algorithmic_bytes": algo_bytes}
        if measured:
            roofline["traffic_source"] = measured["source"]
            roofline["traffic_gbs"] = round(measured["bytes"] / (proj_ms * 1e-3) / 1e9, 1)
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
                   "profile": wl.profile[0], "phase_delay": wl.params.get("phase_delay", "off")},
        "scatterer_evals_per_s": round(evals / (ms * 1e-3), 1),
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        "pcie_inclusive_lines_per_s": round(L / (host_ms * 1e-3), 1),
        "roofline": roofline,
    }
    if not args.no_configs:
        line["configs"] = run_configs()
    if not args.no_batch:
        line["batched_64_jobs"] = run_batched(64, args.batch_scale)
    if not args.no_serve:
        line["serving"] = run_serving(args.serve_seconds, args.serve_rate)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
