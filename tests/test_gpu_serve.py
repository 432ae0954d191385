"""Serving loop: resident phantoms, memory-budget admission, continuous batching.

Every request's IQ is checked against a single simulator configured with the
same phantom and scan (every timestamp t) -- equal up to float summation order,
rel L2 < 1e-6 -- and sampled requests against the FP64 oracle at the parity bar
(rel L2 <= 1e-4).  The "alone vs in the mix" test runs the same requests alone
and inside bursts of other requests."""
import threading

import numpy as np
import pytest

from backends import Oracle, rel_l2
from openbcsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_sim(wl):
    from openbcsim_amd.pyrfsim import RfSimulator
    s = RfSimulator("gpu")
    wl.apply(s)
    return s


@pytest.fixture(scope="module")
def setup():
    from openbcsim_amd.serve import Server
    phantoms, geometries, pairs = W.serving_mix(seed=5, scale=0.02)
    sims = [make_sim(wl) for wl in phantoms]
    srv = Server()
    ids = [srv.add_phantom(s) for s in sims]
    gids = [srv.add_geometry(o, d, l, length) for (o, d, l, length) in geometries]
    assert ids == list(range(len(phantoms))) and gids == list(range(len(geometries)))
    yield srv, phantoms, geometries, pairs, sims
    srv.close()


def single_frame(phantoms, geometries, p, g, t):
    wl = W.at_time(phantoms[p], geometries[g], t)
    s = make_sim(wl)
    f = s.simulate_frame()
    s.close()
    return f


def test_requests_match_single_simulator_and_oracle(setup):
    srv, phantoms, geometries, pairs, _ = setup
    rng = np.random.default_rng(1)
    chosen = [pairs[i] for i in rng.choice(len(pairs), 8, replace=False)]
    reqs = [(p, g, float(rng.uniform(0, 1))) for p, g in chosen]
    handles = [srv.submit(p, g, t) for p, g, t in reqs]
    for (p, g, t), h in zip(reqs, handles):
        iq = srv.wait(h)
        ref = single_frame(phantoms, geometries, p, g, t)
        assert iq.shape == ref.shape
        e = rel_l2(iq, ref)
        o = Oracle()
        W.at_time(phantoms[p], geometries[g], t).apply(o)
        eo = rel_l2(iq, o.simulate_frame())
        print(f"[result] request phantom={p}({phantoms[p].name}) geometry={g} t={t:.3f} "
              f"lines={iq.shape[0]} vs single {e:.2e} vs oracle {eo:.2e} latency {h.latency_ms:.2f} ms", flush=True)
        assert e < 1e-6 and eo <= 1e-4


def test_alone_vs_in_the_mix(setup):
    """A request's result is the same alone and inside bursts of other requests
    (batched into shared frames while earlier frames run)."""
    srv, phantoms, geometries, pairs, _ = setup
    rng = np.random.default_rng(2)
    probes = [(pairs[i][0], pairs[i][1], float(rng.uniform(0, 1))) for i in rng.choice(len(pairs), 4, replace=False)]
    alone = []
    for p, g, t in probes:
        alone.append(srv.wait(srv.submit(p, g, t)).copy())
    srv.set_parameter("reset_stats", 1)
    burst, where = [], []
    for k in range(60):
        if k % 15 == 7:
            where.append(len(burst))
            p, g, t = probes[len(where) - 1]
        else:
            p, g = pairs[int(rng.integers(len(pairs)))]
            t = float(rng.uniform(0, 1))
        burst.append(srv.submit(p, g, t))
    results = [srv.wait(h) for h in burst]
    jobs = srv.stats("frame_jobs")
    print(f"[result] burst of {len(burst)} requests ran in {len(jobs)} frames, max {int(jobs.max())} per frame",
          flush=True)
    assert jobs.max() > 1, "requests that arrive together share frames"
    for (p, g, t), a, i in zip(probes, alone, where):
        e = rel_l2(results[i], a)
        print(f"[result] probe phantom={p} geometry={g}: mix vs alone {e:.2e}", flush=True)
        assert e < 1e-6


def test_pinned_outputs_and_completion_order(setup):
    """Outputs in the server's pinned buffers (alloc_output) receive the IQ by
    DMA with no copy in wait; in a burst mixing sizes every request completes
    with its own frame's result, whatever order the copies land in.  A buffer
    still written by an unreported request cannot be freed."""
    from openbcsim_amd._lib import BcsimError
    srv, phantoms, geometries, pairs, _ = setup
    rng = np.random.default_rng(4)
    work = [(pairs[i][0], pairs[i][1], float(rng.uniform(0, 1))) for i in rng.choice(len(pairs), 10, replace=False)]
    expected = [single_frame(phantoms, geometries, p, g, t) for p, g, t in work]
    outs = [srv.alloc_output(p, g) for p, g, _ in work]
    handles = [srv.submit(p, g, t, out=o) for (p, g, t), o in zip(work, outs)]
    with pytest.raises(BcsimError):
        srv.free_output(outs[-1])  # its request is not reported yet
    for k, h in enumerate(handles):
        iq = srv.wait(h)
        assert iq.__array_interface__["data"][0] == outs[k].__array_interface__["data"][0]
        e = rel_l2(outs[k], expected[k])
        print(f"[result] pinned output request {k} lines={iq.shape[0]}: vs single {e:.2e}, "
              f"latency {h.latency_ms:.2f} ms", flush=True)
        assert e < 1e-6
    copies = srv.stats("copy_ms")
    assert float(copies[-len(work):].max()) < 0.05, "no staging copy for pinned outputs"
    for o in outs:
        srv.free_output(o)
    # pageable and pinned outputs of one request agree
    p, g, t = work[0]
    a = srv.wait(srv.submit(p, g, t)).copy()
    o = srv.alloc_output(p, g)
    b = srv.wait(srv.submit(p, g, t, out=o)).copy()
    srv.free_output(o)
    assert rel_l2(a, b) < 1e-6


def test_requests_served_in_pieces(setup):
    """Requests longer than piece_lines run as line ranges in successive frames
    (at most one piece of a request per frame, kernels of piece frames one after
    another); each request's IQ still equals its single-simulator frame, into
    pageable and into pinned outputs, mixed with unsplit requests."""
    srv, phantoms, geometries, pairs, _ = setup
    rng = np.random.default_rng(6)
    big = [pg for pg in pairs if geometries[pg[1]][0].shape[0] in (256, 512)]
    small = [pg for pg in pairs if geometries[pg[1]][0].shape[0] <= 64]
    work = [big[i] + (float(rng.uniform(0, 1)),) for i in rng.choice(len(big), 4, replace=False)]
    work += [small[i] + (float(rng.uniform(0, 1)),) for i in rng.choice(len(small), 4, replace=False)]
    expected = [single_frame(phantoms, geometries, p, g, t) for p, g, t in work]
    srv.set_parameter("piece_lines", 100)  # 256 lines: 100 + 100 + 56
    try:
        for pinned in (False, True):
            outs = [srv.alloc_output(p, g) if pinned else None for p, g, _ in work]
            srv.set_parameter("reset_stats", 1)
            handles = [srv.submit(p, g, t, out=o) for (p, g, t), o in zip(work, outs)]
            for k, h in enumerate(handles):
                iq = srv.wait(h)
                e = rel_l2(iq, expected[k])
                print(f"[result] pieces pinned={pinned} request {k} lines={iq.shape[0]}: vs single {e:.2e}, "
                      f"latency {h.latency_ms:.2f} ms", flush=True)
                assert iq.shape == expected[k].shape and e < 1e-6
            lines = srv.stats("frame_lines")
            assert lines.sum() == sum(x.shape[0] for x in expected)
            most = max(-(-x.shape[0] // 100) for x in expected)
            assert len(lines) >= most  # one piece of a request per frame
            for o in outs:
                if o is not None:
                    srv.free_output(o)
    finally:
        srv.set_parameter("piece_lines", 1024)
    from openbcsim_amd.pyrfsim import BcsimError
    with pytest.raises(BcsimError, match="piece_lines"):
        srv.set_parameter("piece_lines", -1)
    with pytest.raises(BcsimError, match="frame_contexts"):
        srv.set_parameter("frame_contexts", 5)


def test_concurrent_submitters(setup):
    srv, phantoms, geometries, pairs, _ = setup
    rng = np.random.default_rng(3)
    work = [(pairs[i][0], pairs[i][1], float(rng.uniform(0, 1))) for i in rng.choice(len(pairs), 6, replace=False)]
    expected = [single_frame(phantoms, geometries, p, g, t) for p, g, t in work]
    errors = []

    def client(seed):
        r = np.random.default_rng(seed)
        for _ in range(8):
            k = int(r.integers(len(work)))
            p, g, t = work[k]
            iq = srv.wait(srv.submit(p, g, t))
            e = rel_l2(iq, expected[k])
            if e >= 1e-6:
                errors.append((k, e))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors


def test_memory_budget_admission():
    """With a budget that holds only a few frames' buffers, every request still
    completes correctly (admitted as buffers free up); accounted bytes stay within
    the budget; a request that can never fit fails with an out-of-memory status."""
    from openbcsim_amd.serve import Server
    from openbcsim_amd.pyrfsim import BcsimError
    wl = W.fixed_sector_2d(20000, 16, seed=81)
    sim = make_sim(wl)
    small = W.sector_scan(16, np.radians(60.0), 0.12)
    big = W.sector_scan(512, np.radians(60.0), 0.12)
    ref = single_frame([wl], [small[:3] + (small[4],)], 0, 0, 0.0)
    srv = Server(memory_budget_bytes=1 << 30)
    p = srv.add_phantom(sim)
    resident = srv.stats("resident_bytes")[0]
    g_small = srv.add_geometry(*small[:3], small[4])
    g_big = srv.add_geometry(*big[:3], big[4])
    # room for the phantom plus about three small requests' buffers and outputs
    per_req = 16 * 7800 * 4 + 16384 * 16 + 16 * 7800 * 8 + 16 * 64
    srv.set_parameter("memory_budget_bytes", int(resident + 3.5 * per_req))
    handles = [srv.submit(p, g_small, 0.0) for _ in range(20)]
    for h in handles:
        assert rel_l2(srv.wait(h), ref) < 1e-6
    assert srv.stats("used_bytes")[0] <= srv.stats("budget_bytes")[0]
    assert srv.stats("frame_jobs").max() <= 4
    h = srv.submit(p, g_big, 0.0)
    with pytest.raises(BcsimError, match="memory budget"):
        srv.wait(h)
    srv.close()


def test_server_errors(setup):
    from openbcsim_amd.pyrfsim import BcsimError
    srv = setup[0]
    with pytest.raises(BcsimError, match="unknown phantom"):
        srv.submit(99, 0, 0.0)
    with pytest.raises(BcsimError, match="unknown geometry"):
        srv.submit(0, 99, 0.0)
    with pytest.raises(BcsimError, match="capacity"):
        srv.submit(0, 0, 0.0, out=np.empty(4, np.complex64))
    with pytest.raises(BcsimError, match="unknown request"):
        from openbcsim_amd.serve import Pending
        srv.wait(Pending(123456789, np.empty(1, np.complex64)))
    with pytest.raises(BcsimError, match="illegal parameter"):
        srv.set_parameter("no_such_key", 1)
