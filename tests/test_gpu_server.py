"""Continuous-batching server: requests that arrive while others are on the
GPU join the next step, and every design's result is bit-identical to the
same design evaluated in a standalone batch (integer accumulation)."""
import time

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N

pytestmark = pytest.mark.gpu


def small_requests(sizes, seed=0):
    """Requests of config-2 designs with reduced sampling so the test is quick."""
    rng = np.random.default_rng(seed)
    _, D, S, _ = cpd.generate_designs(2, 0, 256)
    reqs = []
    for n in sizes:
        idx = rng.integers(0, len(D), n)
        designs = [D[i] for i in idx]
        samp = [N.Sampling(int(rng.integers(4, 24)), int(64 * rng.integers(16, 160)), N.MODE_FIXED, 0) for _ in idx]
        reqs.append((designs, samp))
    return reqs


def test_server_mixed_arrivals_match_standalone_batches():
    reqs = small_requests([40, 7, 120, 1, 60])
    ref = [cpd.fitness_batch(d, s)[0] for d, s in reqs]
    with cpd.Server(step_items=512) as srv:
        ids = [srv.submit(*reqs[0]), srv.submit(*reqs[1])]
        time.sleep(0.002)  # the first requests are in flight when the rest arrive
        ids += [srv.submit(*r) for r in reqs[2:]]
        got = [srv.wait(i, 60000) for i in ids]
        stats = srv.stats()
    for (res, status, lat), want in zip(got, ref):
        assert (status == 0).all()
        assert np.array_equal(res, want)
        assert lat > 0
    assert stats["completed"] == len(reqs) and stats["steps"] > 1


def test_server_admission_by_slots():
    # 64 design slots hold two 30-design requests at a time: the others queue
    reqs = small_requests([30, 30, 30, 30], seed=1)
    with cpd.Server(max_designs=64, step_items=256) as srv:
        ids = [srv.submit(*r) for r in reqs]
        got = [srv.wait(i, 60000) for i in ids]
        # a request that can never fit is refused up front
        big_d, big_s = small_requests([65], seed=2)[0]
        with pytest.raises(cpd.CudaError):
            srv.submit(big_d, big_s)
        assert srv.stats()["rejected"] == 1
    for (res, status, _), (d, s) in zip(got, reqs):
        assert np.array_equal(res, cpd.fitness_batch(d, s)[0])


def test_server_reports_invalid_designs_per_request():
    (designs, samp), = small_requests([5], seed=3)
    samp[2] = N.Sampling(0, 1024, N.MODE_FIXED, 0)  # empty sampling
    with cpd.Server() as srv:
        res, status, _ = srv.wait(srv.submit(designs, samp), 60000)
    assert status[2] == N.ERR_INVALID_ARGUMENT and np.isnan(res[2]).all()
    assert (np.delete(status, 2) == 0).all()


def test_server_concurrent_submitters_match_standalone_batches():
    # four threads submit 32 requests of 1..48 designs between them while the
    # server runs; every result equals the standalone batch of its designs
    import threading
    reqs = small_requests(list(np.random.default_rng(7).integers(1, 49, 32)), seed=7)
    ref = [cpd.fitness_batch(d, s)[0] for d, s in reqs]
    got = [None] * len(reqs)
    with cpd.Server(step_items=4096) as srv:
        def worker(w):
            for i in range(w, len(reqs), 4):
                rid = srv.submit(*reqs[i])
                got[i] = srv.wait(rid, 60000)
        threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        stats = srv.stats()
    for (res, status, _), want in zip(got, ref):
        assert (status == 0).all()
        assert np.array_equal(res, want)
    assert stats["completed"] == len(reqs)


def test_two_contexts_in_parallel_threads():
    import threading
    _, D, S, _ = cpd.generate_designs(2, 300, 64)
    samp = [N.Sampling(s.n_wavelengths // 8 + 1, s.n_rays // 16 // 64 * 64 + 64, N.MODE_FIXED, 0) for s in S]
    ref = cpd.fitness_batch(D, samp)[0]
    out = {}

    def run(tag, lo, hi):
        ctx = cpd.Context(0)
        for _ in range(3):
            out[tag] = cpd.fitness_batch(D[lo:hi], samp[lo:hi], ctx=ctx)[0]
        ctx.close()

    threads = [threading.Thread(target=run, args=("a", 0, 32)), threading.Thread(target=run, args=("b", 32, 64))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert np.array_equal(np.vstack([out["a"], out["b"]]), ref)


# ---------------------------------------------------------------------------
# The device-resident loop (resident=True): the same results, bit for bit.

def test_resident_server_matches_standalone_batches():
    reqs = small_requests([40, 7, 120, 1, 60], seed=11)
    ref = [cpd.fitness_batch(d, s)[0] for d, s in reqs]
    with cpd.Server(resident=True) as srv:
        ids = [srv.submit(*reqs[0]), srv.submit(*reqs[1])]
        time.sleep(0.002)  # the first requests are on the GPU when the rest arrive
        ids += [srv.submit(*r) for r in reqs[2:]]
        got = [srv.wait(i, 60000) for i in ids]
        stats = srv.stats()
    for (res, status, lat), want in zip(got, ref):
        assert (status == 0).all()
        assert np.array_equal(res, want)
        assert lat > 0
    assert stats["completed"] == len(reqs) and stats["designs"] == 228 and stats["steps"] >= 1


def test_resident_server_relaunches_after_retiring():
    # with no open request the loop retires; a later request launches it again
    reqs = small_requests([33, 5, 70], seed=12)
    with cpd.Server(resident=True) as srv:
        for r in reqs:
            res, status, _ = srv.wait(srv.submit(*r), 60000)
            assert (status == 0).all() and np.array_equal(res, cpd.fitness_batch(*r)[0])
            time.sleep(0.05)
        assert srv.stats()["steps"] >= 2


def test_resident_server_concurrent_submitters_and_slot_reuse():
    # 64 design slots: requests queue for slots, slots and chunk positions are
    # reused, and four threads submit while the loop runs
    import threading
    reqs = small_requests(list(np.random.default_rng(13).integers(1, 49, 40)), seed=13)
    ref = [cpd.fitness_batch(d, s)[0] for d, s in reqs]
    got = [None] * len(reqs)
    with cpd.Server(resident=True, max_designs=64) as srv:
        def worker(w):
            for i in range(w, len(reqs), 4):
                got[i] = srv.wait(srv.submit(*reqs[i]), 60000)
        threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        stats = srv.stats()
    for (res, status, _), want in zip(got, ref):
        assert (status == 0).all()
        assert np.array_equal(res, want)
    assert stats["completed"] == len(reqs)


def test_resident_server_limits():
    (designs, samp), = small_requests([6], seed=14)
    with cpd.Server(resident=True, max_bins=32, max_rays=4096) as srv:
        big_bins = [d for d in designs if d.bin_count > 32][:1]
        if big_bins:
            with pytest.raises(cpd.CudaError):
                srv.submit(big_bins, [N.Sampling(4, 1024, N.MODE_FIXED, 0)])
        with pytest.raises(cpd.CudaError):
            srv.submit(designs[:1], [N.Sampling(4, 8192, N.MODE_FIXED, 0)])
        small = [d for d in designs if d.bin_count <= 32]
        if small:
            s = [N.Sampling(4, 4096, N.MODE_FIXED, 0)] * len(small)
            res, status, _ = srv.wait(srv.submit(small, s), 60000)
            assert np.array_equal(res, cpd.fitness_batch(small, s)[0])


def test_resident_server_ring_laps():
    # 32 design slots and 64 chunk positions: 240 small requests from three
    # threads wrap the slot pool and the chunk ring many times over; every
    # result equals the standalone batch of its designs
    import threading
    sizes = list(np.random.default_rng(15).integers(1, 21, 240))
    reqs = small_requests(sizes, seed=15)
    reqs = [(d, [N.Sampling(s.n_wavelengths, min(s.n_rays, 2048), N.MODE_FIXED, 0) for s in ss]) for d, ss in reqs]
    got = [None] * len(reqs)
    with cpd.Server(resident=True, max_designs=32) as srv:
        def worker(w):
            for i in range(w, len(reqs), 3):
                got[i] = srv.wait(srv.submit(*reqs[i]), 60000)
        threads = [threading.Thread(target=worker, args=(w,)) for w in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        stats = srv.stats()
    flat_d = [d for r in reqs for d in r[0]]
    flat_s = [s for r in reqs for s in r[1]]
    ref, rst = cpd.fitness_batch(flat_d, flat_s)
    assert (rst == 0).all()
    k = 0
    for (res, status, _), (d, _s) in zip(got, reqs):
        assert (status == 0).all()
        assert np.array_equal(res, ref[k:k + len(d)])
        k += len(d)
    assert stats["completed"] == len(reqs)
