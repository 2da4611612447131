"""The serving loop (msa_srv_*): jobs admitted as they arrive, overlapping on
the device, each job's scores identical alone and in the mix."""
import ctypes
import threading
import time

import numpy as np
import pytest

from msa_amd import _native as N
from msa_amd import workloads as W
from tests import helpers as H


def make_jobs(rng, count, lo_n=4, hi_n=40, lo_len=20, hi_len=700, dirty_every=5):
    jobs = []
    for j in range(count):
        n = int(rng.integers(lo_n, hi_n + 1))
        jobs.append(H.random_db(rng, n, lo_len, hi_len, dirty=(j % dirty_every == 0)))
    return jobs


def test_server_needs_a_gpu_when_none_is_present():
    if N.device_count() > 0:
        pytest.skip("a GPU is present")
    T, p = H.tables()["blosum62"]
    with pytest.raises(N.MsaError):
        N.Server(T, p)


@pytest.mark.gpu
def test_jobs_alone_and_in_the_mix():
    rng = np.random.default_rng(21)
    T, p = H.tables()["blosum62"]
    jobs = make_jobs(rng, 14)
    alone = [N.align(*db, T, p) for db in jobs]
    for k in (0, 3):
        assert np.array_equal(alone[k].astype(np.float32), H.oracle_scores(jobs[k], T, p))
    srv = N.Server(T, p, slots=4, max_length=1024)
    try:
        ids = [srv.submit(*db) for db in jobs]          # a burst: the mix
        results = []
        for jid in ids:
            srv.wait(jid, 60000)
            results.append(srv.fetch(jid))
        for (scores, st), ref in zip(results, alone):
            assert np.array_equal(scores, ref)
            assert st["arrival_us"] <= st["admitted_us"] <= st["finished_us"]
        # Some jobs ran on the device at the same time.
        spans = sorted((st["admitted_us"], st["finished_us"]) for _, st in results)
        overlaps = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
        assert overlaps >= 1
        assert srv.stats()["completed"] == len(jobs)
    finally:
        srv.close()


@pytest.mark.gpu
def test_admission_waits_for_arena_space():
    """An arena that holds about one job at a time still serves every job, in
    arrival order; a job larger than the arena is refused at submission."""
    rng = np.random.default_rng(22)
    T, p = H.tables()["pam250"]
    jobs = make_jobs(rng, 6, lo_n=30, hi_n=40, lo_len=200, hi_len=400, dirty_every=100)
    alone = [N.align(*db, T, p) for db in jobs]
    srv = N.Server(T, p, slots=4, max_length=512, device_bytes=80 << 10, host_bytes=1 << 20)
    try:
        ids = [srv.submit(*db) for db in jobs]
        stats = []
        for jid, ref in zip(ids, alone):
            srv.wait(jid, 60000)
            scores, st = srv.fetch(jid)
            assert np.array_equal(scores, ref)
            stats.append(st)
        admitted = [st["admitted_us"] for st in stats]
        assert admitted == sorted(admitted)
        big = H.random_db(rng, 120, 400, 500)
        with pytest.raises(N.MsaError) as e:
            srv.submit(*big)
        assert e.value.status == N.MSA_ERR_TOO_LARGE
    finally:
        srv.close()


@pytest.mark.gpu
def test_poisson_stream_with_concurrent_submitter():
    """Jobs submitted from another thread at Poisson times while results are
    collected: every job completes and matches the batch entry point."""
    tables = H.tables()
    jobs = W.serving_jobs(tables, 10, seed=5, min_seqs=8, max_seqs=48, lo=100, hi=600)
    T, p = tables["blosum62"]
    alone = [N.align(j.blocks, j.offsets, j.sizes, T, p) for j in jobs]
    arrivals = W.poisson_arrivals(200.0, len(jobs), seed=50)
    srv = N.Server(T, p, slots=4, max_length=2048)
    ids = [None] * len(jobs)
    try:
        t0 = time.time()

        def submitter():
            for k, j in enumerate(jobs):
                delay = arrivals[k] - (time.time() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(j.blocks, j.offsets, j.sizes)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for k, jid in enumerate(ids):
            srv.wait(jid, 60000)
            scores, _ = srv.fetch(jid)
            assert np.array_equal(scores, alone[k])
    finally:
        srv.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rate", [10.0, 500.0])
def test_serving_workload_poisson(rate):
    """The serving workload itself: jobs of 8-256 protein sequences (count
    uniform), L ~ U[100, 2000], BLOSUM62 background residues, blosum62, seed
    5; Poisson arrivals at `rate` jobs/s (seed 50), 24 jobs submitted from
    another thread while results are collected.  Every job equals the same
    job scored alone (msa_nw_align), a sample of its pairs equals the
    oracle, and at 500 jobs/s jobs run on the device at the same time."""
    tables = H.tables()
    jobs = W.serving_jobs(tables, 24, seed=5)
    assert min(j.nseq for j in jobs) >= 8 and max(j.nseq for j in jobs) <= 256
    assert min(j.lengths.min() for j in jobs) >= 100 and max(j.lengths.max() for j in jobs) <= 2000
    T, p = tables["blosum62"]
    alone = [N.align(j.blocks, j.offsets, j.sizes, T, p) for j in jobs]
    rng = np.random.default_rng(7)
    for j, ref in zip(jobs, alone):
        idx = rng.choice(j.npairs, 3, replace=False)
        pairs = np.array([W.pair_of(int(c)) for c in idx], dtype=np.int32)
        exp = H.oracle_scores((j.blocks, j.offsets, j.sizes), T, p, pairs=pairs)
        assert np.array_equal(ref[idx].astype(np.float32), exp)
    arrivals = W.poisson_arrivals(rate, len(jobs), seed=50)
    srv = N.Server(T, p, slots=4, max_length=2048, device_bytes=4 << 30, host_bytes=512 << 20)
    ids = [None] * len(jobs)
    errors = []
    try:
        t0 = time.time()

        def submitter():
            try:
                for k, j in enumerate(jobs):
                    delay = arrivals[k] - (time.time() - t0)
                    if delay > 0:
                        time.sleep(delay)
                    ids[k] = srv.submit(j.blocks, j.offsets, j.sizes)
            except Exception as e:      # surfaced below
                errors.append(e)

        th = threading.Thread(target=submitter)
        th.start()
        stats = []
        for k in range(len(jobs)):
            while ids[k] is None and th.is_alive():
                time.sleep(0.0005)
            assert ids[k] is not None, errors
            srv.wait(ids[k], 60000)
            scores, st = srv.fetch(ids[k])
            assert np.array_equal(scores, alone[k]), "job %d differs from the job scored alone" % k
            assert st["pairs"] == jobs[k].npairs
            stats.append(st)
        th.join()
        assert not errors
        spans = sorted((st["admitted_us"], st["finished_us"]) for st in stats)
        overlaps = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
        if rate >= 500:
            assert overlaps >= 1
        assert srv.stats()["completed"] == len(jobs)
    finally:
        srv.close()


def _concat(dbs):
    """One store holding the sequences of several stores, in order."""
    blocks, offsets, sizes, base = [], [], [], 0
    for b, o, s in dbs:
        blocks.append(np.asarray(b, dtype=np.uint16))
        offsets.append(np.asarray(o, dtype=np.int64) + base)
        sizes.append(np.asarray(s, dtype=np.int32))
        base += len(b)
    return np.concatenate(blocks), np.concatenate(offsets), np.concatenate(sizes)


def _select(db, idx):
    b, o, s = db
    return b, np.asarray(o)[idx], np.asarray(s)[idx]


@pytest.mark.gpu
def test_resident_store_jobs_upload_no_sequences():
    """Sequences put once into the server's store and named by id: a resident
    job uploads only its work lists (no sequence bytes), its scores equal the
    same sequences scored alone, also with ids drawn across two store_add
    calls in any order and mixed with uploading jobs; the store's sets are
    held while jobs read them and removed ids are refused."""
    rng = np.random.default_rng(31)
    T, p = H.tables()["blosum62"]
    a = H.random_db(rng, 40, 50, 600, dirty=True)
    b = H.random_db(rng, 30, 100, 900)
    both = _concat([a, b])
    srv = N.Server(T, p, slots=4, max_length=1024, store_bytes=8 << 20)
    try:
        first_a = srv.store_add(*a)
        first_b = srv.store_add(*b)
        assert first_b == first_a + 40
        st = srv.stats()
        assert st["store_sequences"] == 70 and st["store_bytes_used"] >= 2 * (len(a[0]) + len(b[0]))
        picks = [np.arange(40), np.arange(40, 70), rng.permutation(70)[:25], rng.permutation(70)[:50],
                 np.array([69, 0, 41, 3, 40, 2])]
        jobs = []
        for idx in picks:
            jobs.append(("resident", idx, srv.submit_resident(first_a + idx)))
            jobs.append(("upload", idx, srv.submit(*_select(both, idx))))
        got = {}
        for kind, idx, jid in jobs:
            srv.wait(jid, 60000)
            got[(kind, tuple(idx))] = srv.fetch(jid)
        for kind, idx, _ in jobs:
            scores, stats = got[(kind, tuple(idx))]
            ref = N.align(*_select(both, idx), T, p)
            assert np.array_equal(scores, ref), (kind, idx[:5])
            if kind == "resident":
                assert stats["sequence_bytes"] == 0
                up = got[("upload", tuple(idx))][1]
                assert up["sequence_bytes"] > 0
                assert stats["upload_bytes"] <= up["upload_bytes"] - up["sequence_bytes"] + 256
        # A set that a queued or running job reads cannot go; once the job
        # has ended it can.
        jid = srv.submit_resident(first_b + np.arange(30))
        try:
            srv.store_remove(first_b)
            assert srv.done(jid)           # refused only while the job runs
        except N.MsaError as e:
            assert e.status == N.MSA_ERR_BUSY
            srv.wait(jid, 60000)
            srv.store_remove(first_b)
        assert np.array_equal(srv.fetch(jid)[0], N.align(*b, T, p))
        srv.store_remove(first_a)
        assert srv.stats()["store_sequences"] == 0 and srv.stats()["store_bytes_used"] == 0
        with pytest.raises(N.MsaError) as e:
            srv.submit_resident(np.array([first_a, first_a + 1]))
        assert e.value.status == N.MSA_ERR_NOT_FOUND
        with pytest.raises(N.MsaError) as e:
            srv.store_remove(first_a)
        assert e.value.status == N.MSA_ERR_NOT_FOUND
    finally:
        srv.close()


@pytest.mark.gpu
def test_scores_stay_in_the_arena_until_fetched():
    """A finished job's scores hold their arena block until fetched, to the
    host or to device memory on a stream; both give the alone scores."""
    rng = np.random.default_rng(32)
    T, p = H.tables()["pam250"]
    jobs = make_jobs(rng, 4, lo_n=20, hi_n=30, lo_len=100, hi_len=300, dirty_every=100)
    alone = [N.align(*db, T, p) for db in jobs]
    srv = N.Server(T, p, slots=2, max_length=512)
    try:
        ids = [srv.submit(*db) for db in jobs]
        for jid in ids:
            srv.wait(jid, 60000)
        held = srv.stats()["device_bytes_used"]
        assert held >= sum(4 * len(r) for r in alone)
        scores, st = srv.fetch(ids[0])
        assert np.array_equal(scores, alone[0]) and st["fetch_us"] > 0
        assert srv.stats()["device_bytes_used"] < held
        L = N.lib()
        stream = ctypes.c_void_p()
        N.check(L.msa_stream_create(ctypes.byref(stream)))
        for jid, ref in zip(ids[1:], alone[1:]):
            buf = N.DeviceBuffer(4 * len(ref))
            srv.fetch_device(jid, buf.ptr, len(ref), stream.value)
            N.check(L.msa_stream_sync(stream.value))
            assert np.array_equal(buf.download(np.int32, len(ref)), ref)
            buf.close()
        deadline = time.time() + 10
        while srv.stats()["device_bytes_used"] and time.time() < deadline:
            time.sleep(0.01)
        assert srv.stats()["device_bytes_used"] == 0
        N.check(L.msa_stream_destroy(stream.value))
    finally:
        srv.close()
