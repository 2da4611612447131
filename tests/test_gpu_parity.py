"""GPU parity: every kernel, through the C ABI, against the oracle.

Small cases are compared score by score with the CPU oracle (the C
restatement of sequential.cpp:29-95, itself pinned in test_oracle.py); the
full-size workloads are checked through properties that do not depend on a
CPU pass over all pairs: batched == plain on every pair, a random oracle
sample, and symmetry.  The bar is bit-exact int32 == the reference's float.
"""
import ctypes
import json
import pathlib
import re
import zlib

import numpy as np
import pytest

from msa_amd import _native as N
from msa_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu

KERNELS = {"auto": N.KERNEL_AUTO, "plain": N.KERNEL_PLAIN, "batched": N.KERNEL_BATCHED, "long": N.KERNEL_LONG}


@pytest.fixture(scope="module", autouse=True)
def device():
    if N.device_count() < 1:
        pytest.skip("no GPU")
    N.check(N.lib().msa_device_select(0))


def gpu_scores(db, T, p, pairs=None, kernel="auto", **kw):
    return N.align(*db, T, p, pairs=pairs, opts=N.options(KERNELS[kernel], **kw))


def assert_oracle(db, T, p, pairs=None, kernel="auto", **kw):
    got = gpu_scores(db, T, p, pairs=pairs, kernel=kernel, **kw)
    exp = H.oracle_scores(db, T, p, pairs=pairs)
    bad = np.nonzero(got.astype(np.float32) != exp)[0]
    assert len(bad) == 0, "%d/%d differ, first %s: gpu %s oracle %s" % (
        len(bad), len(exp), bad[:5], got[bad[:5]], exp[bad[:5]])
    return got


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("key", sorted(H.golden_scores()))
def test_reference_fixtures_golden(key, kernel):
    """test/pairwise.py:13-58: mock1..3 x {default, blosum62, pam250}."""
    fixture, table = key.split("/")
    _, db = H.fixture_db(fixture)
    T, p = H.tables()[table]
    got = gpu_scores(db, T, p, kernel=kernel)
    assert [int(v) for v in got] == H.golden_scores()[key]["scores"]


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("table", ["default", "blosum62", "blosum45", "blosum50", "blosum80", "blosum90", "pam250"])
def test_random_clean(table, kernel):
    rng = np.random.default_rng(zlib.crc32((table + kernel).encode()))
    T, p = H.tables()[table]
    assert_oracle(H.random_db(rng, 14, 0, 90), T, p, kernel=kernel)


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_padding_codes(seed, kernel):
    """Internal padding codes: rows stop at the first, columns copy left."""
    rng = np.random.default_rng(seed)
    T, p = H.tables()["blosum62"]
    assert_oracle(H.random_db(rng, 16, 0, 70, dirty=True), T, p, kernel=kernel)


EDGE_STRINGS = [
    "", "A", "AC", "ACG", "ACGT", "*", "**", "***", "*A", "A*", "AC*GT", "ACGT*", "*ACGT", "A*C*G*T",
    "NNNNXXBJZOU", "acgtMKV", "W" * 65, "GATTACA" * 10, "X\rY", "1234", "MK*" * 30,
]


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("table", ["blosum62", "pam250", "default"])
def test_edge_sequences(table, kernel):
    T, p = H.tables()[table]
    assert_oracle(H.strings_db(EDGE_STRINGS), T, p, kernel=kernel)


def test_asymmetric_table_keeps_reference_orientation():
    """Equal block counts put pair.second on the rows (sequential.cpp:87-91);
    with a non-symmetric table the transposed profile must be used."""
    rng = np.random.default_rng(9)
    T, p = H.tables()["blosum62"]
    Ta = T.copy()
    Ta[0, 1], Ta[1, 0] = 5, -3
    Ta[4, 2], Ta[17, 18] = -2, 6
    db = H.random_db(rng, 24, 1, 120)
    for kernel in KERNELS:
        assert_oracle(db, Ta, p, kernel=kernel)
    db = H.strings_db(["ACGTAC", "CATGCA", "ACG", "CAT"])
    for kernel in KERNELS:
        assert_oracle(db, Ta, p, kernel=kernel)


@pytest.mark.parametrize("lo,hi", [(1, 130), (120, 400), (380, 700), (500, 1100), (900, 2100)])
def test_batched_strip_shapes(lo, hi):
    """Queries of many first-strip heights (2..32 rows per lane) and of one to
    five strips, which exercise the strip boundary line in global memory."""
    rng = np.random.default_rng(lo + hi)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 10, lo, hi)
    assert_oracle(db, T, p, kernel="batched")


@pytest.mark.parametrize("lo,hi,alphabet", [(1, 300, 24), (400, 3000, 24), (2000, 9000, 4), (100, 5000, 4),
                                             (5000, 9000, 20)])
def test_long_kernel_shapes(lo, hi, alphabet):
    """The multi-wave pipelined kernel: strips of 64R rows (R from the
    compact alphabet: 32 for DNA, 8 for protein), more strips than waves,
    boundary rows handed down through global memory."""
    rng = np.random.default_rng(lo * 7 + hi + alphabet)
    T, p = H.tables()["blosum62" if alphabet > 4 else "pam250"]
    db = H.random_db(rng, 7, lo, hi, alphabet=alphabet)
    assert_oracle(db, T, p, kernel="long")


@pytest.mark.parametrize("alphabet", [4, 12, 24])
def test_long_kernel_every_first_strip_height(alphabet):
    """The long kernel's first strip takes the rows the full strips leave, in
    steps of 8 rows per lane: query lengths in every step's range of residues
    (one strip and two), for the 32-, 16- and 8-row strips that 4-, 12- and
    24-code alphabets give; every pair against the oracle."""
    rng = np.random.default_rng(5 + alphabet)
    T, p = H.tables()["blosum62"]
    rows = 32 if alphabet <= 7 else (16 if alphabet <= 15 else 8)
    lengths = []
    for r0 in range(8, rows + 1, 8):
        rem = 64 * (r0 - 8) + 1 + int(rng.integers(0, 64 * 8))
        lengths += [rem, rem + 64 * rows]
    codes = [rng.integers(0, alphabet, n).astype(np.uint8) for n in lengths]
    assert_oracle(W.pack(codes), T, p, kernel="long")


def test_long_kernel_asymmetric_and_padding_routing():
    rng = np.random.default_rng(77)
    T, p = H.tables()["blosum62"]
    Ta = T.copy()
    Ta[0, 1], Ta[1, 0], Ta[2, 3] = 5, -3, 7
    db = H.random_db(rng, 8, 600, 2600, dirty=True, alphabet=4)
    assert_oracle(db, Ta, p, kernel="long")


def test_plain_multi_strip_with_padding():
    rng = np.random.default_rng(4)
    T, p = H.tables()["pam250"]
    db = H.random_db(rng, 8, 500, 2100, dirty=True)
    assert_oracle(db, T, p, kernel="plain")


def test_explicit_pair_list():
    rng = np.random.default_rng(5)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 30, 0, 200, dirty=True)
    pairs = rng.integers(0, 30, (300, 2)).astype(np.int32)
    for kernel in KERNELS:
        assert_oracle(db, T, p, pairs=pairs, kernel=kernel)


def test_uniform_small_config_matches_oracle():
    """16 protein sequences, L ~ U[100,300], blosum62: all 120 pairs."""
    w = W.uniform_small(H.tables())
    for kernel in KERNELS:
        assert_oracle((w.blocks, w.offsets, w.sizes), w.table, w.penalty, kernel=kernel)


def test_graph_replays_match_eager():
    rng = np.random.default_rng(6)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 40, 10, 900, dirty=True)
    n = 40 * 39 // 2
    eager = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=False))
    graph = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=True))
    out_e = N.DeviceBuffer(4 * n)
    out_g = N.DeviceBuffer(4 * n)
    eager.execute(out_e.ptr)
    N.check(N.lib().msa_device_sync())
    ref = out_e.download(np.int32, n)
    assert np.array_equal(ref.astype(np.float32), H.oracle_scores(db, T, p))
    for _ in range(4):
        N.check(N.lib().msa_memset(out_g.ptr, 0xAB, 4 * n))
        graph.execute(out_g.ptr)
        N.check(N.lib().msa_device_sync())
        assert np.array_equal(out_g.download(np.int32, n), ref)
        eager.execute(out_e.ptr)
        N.check(N.lib().msa_device_sync())


def test_caller_captured_graph_of_a_plan():
    """msa_graph_*: the caller captures msa_nw_plan_execute of a plan built
    with use_graph = 0 on its own stream and replays the graph; every replay,
    with eager executions and synchronisations between, equals the eager
    scores (plain, batched and long pairs in one plan)."""
    lib = N.lib()
    rng = np.random.default_rng(8)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 30, 5, 1400, dirty=True)
    n = 30 * 29 // 2
    stream = ctypes.c_void_p()
    N.check(lib.msa_stream_create(ctypes.byref(stream)))
    plan = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=False, long_threshold=1000))
    out_e, out_g = N.DeviceBuffer(4 * n), N.DeviceBuffer(4 * n)
    graph = ctypes.c_void_p()
    try:
        info = plan.info()
        assert info.plain_pairs > 0 and info.batched_pairs > 0 and info.long_pairs > 0
        plan.execute(out_e.ptr, stream.value)
        N.check(lib.msa_stream_sync(stream))
        ref = out_e.download(np.int32, n)
        assert np.array_equal(ref.astype(np.float32), H.oracle_scores(db, T, p))
        N.check(lib.msa_graph_capture_begin(stream))
        plan.execute(out_g.ptr, stream.value)
        N.check(lib.msa_graph_capture_end(stream, ctypes.byref(graph)))
        for _ in range(4):
            N.check(lib.msa_memset(out_g.ptr, 0xAB, 4 * n))
            N.check(lib.msa_graph_launch(graph, stream))
            N.check(lib.msa_stream_sync(stream))
            assert np.array_equal(out_g.download(np.int32, n), ref)
            plan.execute(out_e.ptr, stream.value)
            N.check(lib.msa_stream_sync(stream))
            assert np.array_equal(out_e.download(np.int32, n), ref)
    finally:
        N.check(lib.msa_graph_destroy(graph))
        plan.close()
        out_e.close()
        out_g.close()
        N.check(lib.msa_stream_destroy(stream))


def test_plain_entry_on_device_memory():
    """msa_nw_plain: caller-owned device buffers, all-pairs and explicit list."""
    lib = N.lib()
    rng = np.random.default_rng(8)
    T, p = H.tables()["blosum62"]
    blocks, offsets, sizes = H.random_db(rng, 20, 0, 700, dirty=True)
    n = 20
    d_blocks = N.DeviceBuffer(blocks.nbytes)
    d_blocks.upload(blocks)
    d_off = N.DeviceBuffer(offsets.nbytes)
    d_off.upload(offsets)
    d_size = N.DeviceBuffer(sizes.nbytes)
    d_size.upload(sizes)
    npairs = n * (n - 1) // 2
    d_out = N.DeviceBuffer(4 * npairs)
    table = np.ascontiguousarray(T, dtype=np.int32)
    N.check(lib.msa_nw_plain(d_blocks.ptr, d_off.ptr, d_size.ptr, n, int(sizes.max()), None, npairs,
                             table.ctypes.data, p, d_out.ptr, None))
    N.check(lib.msa_device_sync())
    exp = H.oracle_scores((blocks, offsets, sizes), T, p)
    assert np.array_equal(d_out.download(np.int32, npairs).astype(np.float32), exp)
    pairs = rng.integers(0, n, (77, 2)).astype(np.int32)
    d_pairs = N.DeviceBuffer(pairs.nbytes)
    d_pairs.upload(pairs)
    N.check(lib.msa_nw_plain(d_blocks.ptr, d_off.ptr, d_size.ptr, n, int(sizes.max()), d_pairs.ptr, len(pairs),
                             table.ctypes.data, p, d_out.ptr, None))
    N.check(lib.msa_device_sync())
    exp = H.oracle_scores((blocks, offsets, sizes), T, p, pairs=pairs)
    assert np.array_equal(d_out.download(np.int32, len(pairs)).astype(np.float32), exp)


def _full_size_check(w, sample=200, seed=0):
    db = (w.blocks, w.offsets, w.sizes)
    batched = gpu_scores(db, w.table, w.penalty, kernel="auto")
    plain = gpu_scores(db, w.table, w.penalty, kernel="plain")
    assert np.array_equal(batched, plain)
    rng = np.random.default_rng(seed)
    idx = rng.choice(w.npairs, min(sample, w.npairs), replace=False)
    pairs = np.array([divmod_pair(int(c)) for c in idx], dtype=np.int32)
    exp = H.oracle_scores(db, w.table, w.penalty, pairs=pairs, workers=16)
    assert np.array_equal(batched[idx].astype(np.float32), exp)
    return batched


def divmod_pair(c):
    i = int((1 + np.sqrt(1 + 8 * c)) // 2)
    while i * (i - 1) // 2 > c:
        i -= 1
    while (i + 1) * i // 2 <= c:
        i += 1
    return i, c - i * (i - 1) // 2


def test_full_uniform_batch_512():
    """512 sequences, L ~ U[800,1200]: all 130,816 pairs, batched == plain."""
    _full_size_check(W.uniform_batch(H.tables()))


def test_full_uneven_batch_2000_graph_replays():
    """2000 sequences, log-normal lengths: all 1,999,000 pairs scored by the
    hipGraph-captured plan the bench times (the first execution captures and
    launches, the next ones replay the instantiated graph), each replay equal
    to the plain kernel on every pair, and a random sample equal to the oracle."""
    w = W.uneven_batch(H.tables())
    db = (w.blocks, w.offsets, w.sizes)
    plain = gpu_scores(db, w.table, w.penalty, kernel="plain")
    plan = N.Plan(*db, w.table, w.penalty, opts=N.options(N.KERNEL_AUTO, use_graph=True))
    out = N.DeviceBuffer(4 * w.npairs)
    try:
        for replay in range(3):
            N.check(N.lib().msa_memset(out.ptr, 0xAB, 4 * w.npairs))
            plan.execute(out.ptr)
            plan.status()
            got = out.download(np.int32, w.npairs)
            assert np.array_equal(got, plain), "replay %d differs from the plain kernel" % replay
    finally:
        plan.close()
        out.close()
    rng = np.random.default_rng(0)
    idx = rng.choice(w.npairs, 300, replace=False)
    pairs = np.array([divmod_pair(int(c)) for c in idx], dtype=np.int32)
    exp = H.oracle_scores(db, w.table, w.penalty, pairs=pairs, workers=16)
    assert np.array_equal(plain[idx].astype(np.float32), exp)


def test_long_dna_sample():
    """64 DNA sequences of 30,000 (+5/-4, gap 8): a pair subset vs the oracle."""
    w = W.long_dna()
    rng = np.random.default_rng(1)
    pairs = rng.choice(64, (12, 2)).astype(np.int32)
    assert_oracle((w.blocks, w.offsets, w.sizes), w.table, w.penalty, pairs=pairs, kernel="auto")


def test_long_dna_full_scale():
    """All 2,016 pairs of the 64 x 30,000 DNA set through the long kernel (the
    auto plan routes every pair there): many rounds of the persistent queue and
    its last, partial round.  Every pair equals the plain kernel, 8 pairs equal
    the oracle, and the kernel's bounded wait never gave up."""
    w = W.long_dna()
    db = (w.blocks, w.offsets, w.sizes)
    plan = N.Plan(*db, w.table, w.penalty, opts=N.options(N.KERNEL_AUTO, use_graph=True))
    out = N.DeviceBuffer(4 * w.npairs)
    try:
        info = plan.info()
        assert info.long_pairs == w.npairs == 2016
        plan.execute(out.ptr)
        plan.status()
        got = out.download(np.int32, w.npairs)
    finally:
        plan.close()
        out.close()
    plain = gpu_scores(db, w.table, w.penalty, kernel="plain")
    assert np.array_equal(got, plain)
    idx = np.random.default_rng(2).choice(w.npairs, 8, replace=False)
    pairs = np.array([divmod_pair(int(c)) for c in idx], dtype=np.int32)
    exp = H.oracle_scores(db, w.table, w.penalty, pairs=pairs, workers=8)
    assert np.array_equal(got[idx].astype(np.float32), exp)


def test_plan_executions_on_two_streams_back_to_back():
    """One plan (plain, batched and long pairs; eager and graph forms)
    executed on two streams with no synchronisation between: the plan
    serializes its executions (they share its scratch), so both outputs equal
    an eager run every time."""
    lib = N.lib()
    rng = np.random.default_rng(31)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 36, 5, 1500, dirty=True)
    n = 36 * 35 // 2
    ref = gpu_scores(db, T, p, long_threshold=1000)
    assert np.array_equal(ref.astype(np.float32), H.oracle_scores(db, T, p))
    streams = [ctypes.c_void_p(), ctypes.c_void_p()]
    for s in streams:
        N.check(lib.msa_stream_create(ctypes.byref(s)))
    outs = [N.DeviceBuffer(4 * n), N.DeviceBuffer(4 * n)]
    try:
        for use_graph in (False, True):
            plan = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=use_graph, long_threshold=1000))
            info = plan.info()
            assert info.plain_pairs > 0 and info.batched_pairs > 0 and info.long_pairs > 0
            for _ in range(3):
                for o in outs:
                    N.check(lib.msa_memset(o.ptr, 0xAB, 4 * n))
                plan.execute(outs[0].ptr, streams[0].value)
                plan.execute(outs[1].ptr, streams[1].value)
                plan.execute(outs[0].ptr, streams[0].value)
                for s in streams:
                    N.check(lib.msa_stream_sync(s))
                plan.status(streams[0].value)
                for o in outs:
                    assert np.array_equal(o.download(np.int32, n), ref)
            plan.close()
    finally:
        for o in outs:
            o.close()
        for s in streams:
            N.check(lib.msa_stream_destroy(s))


def test_plain_entry_refuses_stream_capture():
    """msa_nw_plain grows and frees a workspace of its own, so it refuses to be
    captured into a caller's graph (MSA_ERR_INVALID_ARGUMENT) and leaves the
    capture usable."""
    lib = N.lib()
    rng = np.random.default_rng(32)
    T, p = H.tables()["blosum62"]
    blocks, offsets, sizes = H.random_db(rng, 6, 1, 100)
    d_blocks, d_off, d_size = N.DeviceBuffer(blocks.nbytes), N.DeviceBuffer(offsets.nbytes), N.DeviceBuffer(sizes.nbytes)
    d_blocks.upload(blocks)
    d_off.upload(offsets)
    d_size.upload(sizes)
    d_out = N.DeviceBuffer(4 * 15)
    table = np.ascontiguousarray(T, dtype=np.int32)
    stream = ctypes.c_void_p()
    N.check(lib.msa_stream_create(ctypes.byref(stream)))
    plan = N.Plan(blocks, offsets, sizes, T, p)
    graph = ctypes.c_void_p()
    try:
        N.check(lib.msa_graph_capture_begin(stream))
        rc = lib.msa_nw_plain(d_blocks.ptr, d_off.ptr, d_size.ptr, 6, int(sizes.max()), None, 15,
                              table.ctypes.data, p, d_out.ptr, stream)
        assert rc == N.MSA_ERR_INVALID_ARGUMENT
        plan.execute(d_out.ptr, stream.value)
        N.check(lib.msa_graph_capture_end(stream, ctypes.byref(graph)))
        N.check(lib.msa_graph_launch(graph, stream))
        N.check(lib.msa_stream_sync(stream))
        exp = H.oracle_scores((blocks, offsets, sizes), T, p)
        assert np.array_equal(d_out.download(np.int32, 15).astype(np.float32), exp)
    finally:
        N.check(lib.msa_graph_destroy(graph))
        plan.close()
        N.check(lib.msa_stream_destroy(stream))


def test_gap_code_rejected_on_device_path_too():
    T, p = H.tables()["blosum62"]
    with pytest.raises(N.MsaError) as e:
        gpu_scores(H.strings_db(["AC-GT", "ACGT"]), T, p)
    assert e.value.status == N.MSA_ERR_GAP_CODE


def test_symmetry_of_scores():
    """Symmetric tables: swapping the pair members never changes a score."""
    rng = np.random.default_rng(10)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 25, 0, 300)
    pairs = np.array([(i, j) for i in range(25) for j in range(25)], dtype=np.int32)
    got = gpu_scores(db, T, p, pairs=pairs).reshape(25, 25)
    assert (got == got.T).all()


TILES = [(16, 32), (16, 16), (32, 16)]


@pytest.mark.parametrize("lanes,rows", TILES)
@pytest.mark.parametrize("lo,hi", [(1, 140), (100, 520), (250, 1100), (900, 2100)])
def test_batched_tile_shapes(lanes, rows, lo, hi):
    """Every tile shape of the batched kernel, forced for all pairs: first
    strips of many heights (2..rows rows per lane) and queries of one to
    several strips (strip boundary line in global memory)."""
    rng = np.random.default_rng(lanes * 1000 + rows * 100 + lo + hi)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 12, lo, hi)
    assert_oracle(db, T, p, kernel="batched", lanes=lanes, rows=rows)


def first_strip_row_step():
    """kRowStep of msa_amd/csrc/nw_planner.hpp: the rows-per-lane granularity
    of a query's first strip in the batched kernel."""
    text = (pathlib.Path(__file__).resolve().parents[1] / "msa_amd" / "csrc" / "nw_planner.hpp").read_text()
    return int(re.search(r"constexpr int kRowStep = (\d+);", text).group(1))


@pytest.mark.parametrize("lanes,rows", TILES)
def test_batched_every_first_strip_height(lanes, rows):
    """Every first-strip height (kRowStep, 2 kRowStep, .., rows rows per lane,
    the step read from the planner's header) under every tile shape, for
    queries of one strip and of two: query lengths are picked in each
    height's range of residues, and every pair is scored."""
    rng = np.random.default_rng(7 * lanes + rows)
    T, p = H.tables()["blosum62"]
    row_step = first_strip_row_step()
    step = row_step * lanes
    lengths = []
    for j in range(1, rows // row_step + 1):
        rem = step * (j - 1) + 1 + int(rng.integers(0, step))
        lengths += [rem, rem + lanes * rows]
    codes = [rng.integers(0, 24, n).astype(np.uint8) for n in lengths]
    assert_oracle(W.pack(codes), T, p, kernel="batched", lanes=lanes, rows=rows)


@pytest.mark.parametrize("lanes,rows", TILES)
def test_batched_tile_shapes_asymmetric(lanes, rows):
    """Transposed profiles under every tile shape."""
    rng = np.random.default_rng(lanes + rows)
    T, p = H.tables()["blosum62"]
    Ta = T.copy()
    Ta[0, 1], Ta[1, 0], Ta[5, 9], Ta[20, 3] = 5, -3, 4, -4
    db = H.random_db(rng, 16, 1, 700)
    assert_oracle(db, Ta, p, kernel="batched", lanes=lanes, rows=rows)


def test_plan_mixing_every_tile_shape(tmp_path):
    """A tuning table that gives each query-length bin another tile shape: the
    plan launches one kernel per shape (each with its own queue), and every
    score, under an asymmetric table too, matches the oracle."""
    rng = np.random.default_rng(33)
    T, p = H.tables()["blosum62"]
    Ta = T.copy()
    Ta[0, 1], Ta[1, 0], Ta[7, 2], Ta[15, 11] = 6, -2, 3, -4
    path = str(tmp_path / "mixed.json")
    with open(path, "w") as f:
        json.dump({"format": "msa-nw-tuning/1", "long_threshold": 8192,
                   "bins": [{"min_query": 0, "lanes": 16, "rows": 16},
                            {"min_query": 200, "lanes": 32, "rows": 16},
                            {"min_query": 600, "lanes": 16, "rows": 32}]}, f)
    db = H.random_db(rng, 40, 1, 1500)
    for table in (T, Ta):
        assert_oracle(db, table, p, kernel="auto", tuning_file=path)
    plan = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, tuning_file=path))
    try:
        info = plan.info()
        assert info.bins == 3 and info.plain_pairs == 0
        assert info.launches == 6          # queue reset + kernel per shape
    finally:
        plan.close()


def test_unknown_tile_shape_rejected():
    T, p = H.tables()["blosum62"]
    db = H.strings_db(["ACGT", "ACG"])
    for lanes, rows in [(64, 16), (16, 8), (32, 32), (8, 32)]:
        with pytest.raises(N.MsaError) as e:
            gpu_scores(db, T, p, kernel="batched", lanes=lanes, rows=rows)
        assert e.value.status == N.MSA_ERR_INVALID_ARGUMENT


def test_tuner_writes_a_table_that_plans_read(tmp_path):
    """msa_nw_tune on a small uneven database: coordinate descent over whole
    executions (every other tile shape for each populated length bin, then
    every long threshold up to the longest query), a table that
    msa_nw_options.tuning_file reads, and scores that still match the oracle
    under it."""
    rng = np.random.default_rng(12)
    T, p = H.tables()["blosum62"]
    lengths = np.clip(np.exp(rng.normal(np.log(300), 0.9, 48)), 20, 2500).astype(int)
    codes = [rng.integers(0, 20, int(n)).astype(np.uint8) for n in lengths]
    db = W.pack(codes)
    path = str(tmp_path / "tuning.json")
    table = N.tune(*db, T, p, path=path)
    assert table["format"] == "msa-nw-tuning/1"
    assert table["sample"]["all_pairs"] and table["sample"]["pairs"] == 48 * 47 // 2
    assert table["bins"][0]["min_query"] == 0
    mins = [b["min_query"] for b in table["bins"]]
    assert mins == sorted(set(mins))
    for b in table["bins"]:
        assert (b["lanes"], b["rows"]) in TILES
    populated = [b for b in table["bin_cells"] if b["cells"] > 0]
    assert sum(b["cells"] for b in table["bin_cells"]) == table["sample"]["cells"]
    steps = [m["step"] for m in table["measurements"]]
    assert steps[0] == "start"
    assert len([s for s in steps if s.startswith("bin ")]) == 2 * len(populated)
    tried = [t["long_threshold"] for t in table["threshold_trials"]]
    assert tried == [2 ** 31 - 1] + [th for th in (1024, 2048, 4096, 8192, 16384) if th <= np.sort(lengths)[-2]]
    # A threshold is adopted only when it saves more than 0.5% of the time.
    times = {t["long_threshold"]: t["ms"] for t in table["threshold_trials"]}
    assert table["long_threshold"] in times
    if table["long_threshold"] != 2 ** 31 - 1:
        assert times[table["long_threshold"]] < 0.995 * times[2 ** 31 - 1]
    assert all(m["ms"] > 0 for m in table["measurements"])
    with open(path) as f:
        assert json.load(f)["bins"] == table["bins"]
    assert_oracle(db, T, p, kernel="auto", tuning_file=path)


@pytest.mark.parametrize("case", [1, 2, 3, 4, 5, 6])
def test_reference_benchmark_databases(case):
    """The reference's own benchmark inputs (db/case1..6.fasta: 50-800 viral
    genomes of 2k-10k nt with N and ambiguity letters, 70-column lines and an
    empty line after each record), parsed by the GPU FASTA loader into exactly
    the oracle parser's store, then scored all-pairs with the reference's
    default table: the auto plan (long, batched and plain routing) equals the
    plain kernel on every pair and the oracle on a sample."""
    from oracle import museqa_ref as M
    tables = H.tables()
    w = W.reference_case(case, tables)
    exp_store = M.FastaOracle().parse(W.reference_case_text(case))
    for got, exp in zip((w.blocks, w.offsets, w.sizes), exp_store[:3]):
        assert np.array_equal(got, exp)
    db = (w.blocks, w.offsets, w.sizes)
    auto = gpu_scores(db, w.table, w.penalty, use_graph=True)
    plain = gpu_scores(db, w.table, w.penalty, kernel="plain")
    assert np.array_equal(auto, plain)
    rng = np.random.default_rng(case)
    idx = rng.choice(w.npairs, 6, replace=False)
    pairs = np.array([divmod_pair(int(c)) for c in idx], dtype=np.int32)
    exp = H.oracle_scores(db, w.table, w.penalty, pairs=pairs, workers=6)
    assert np.array_equal(auto[idx].astype(np.float32), exp)


def plan_info(db, T, p, pairs=None, **kw):
    plan = N.Plan(*db, T, p, pairs=pairs, opts=N.options(N.KERNEL_AUTO, **kw))
    try:
        return plan.info()
    finally:
        plan.close()


@pytest.mark.parametrize("lanes,rows", TILES)
def test_stacked_strips_every_tile_shape(lanes, rows):
    """Short queries stacked several to a strip (each below a neutral
    separator row, its rows offset by its own K), under every tile shape and
    an asymmetric table (members of one stack in both orientations): every
    score equals the oracle's and the unstacked plan's, and the plan reports
    stacked pairs."""
    rng = np.random.default_rng(91 + lanes + rows)
    T, p = H.tables()["blosum62"]
    Ta = T.copy()
    Ta[0, 1], Ta[1, 0], Ta[6, 13], Ta[19, 4] = 7, -3, 2, -4
    lengths = list(rng.integers(rows, lanes * rows // 2, 36)) + list(rng.integers(1, 2000, 8))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    for table in (T, Ta):
        got = assert_oracle(db, table, p, lanes=lanes, rows=rows)
        flat = gpu_scores(db, table, p, lanes=lanes, rows=rows, stack=False)
        assert np.array_equal(got, flat)
        info = plan_info(db, table, p, lanes=lanes, rows=rows)
        assert info.stacked_pairs > 0
        assert plan_info(db, table, p, lanes=lanes, rows=rows, stack=False).stacked_pairs == 0


def test_stacked_strips_explicit_list_with_repeats():
    """An explicit pair list naming pairs twice, in both orders and out of
    order, with queries that stack: every listed pair gets its own score
    (a repeated (member, target) is its own union entry), equal to the
    oracle and to the unstacked plan."""
    rng = np.random.default_rng(57)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 30, 40, 300)
    n = len(db[2])
    pairs = [(int(a), int(b)) for a, b in rng.integers(0, n, (400, 2)) if a != b]
    pairs += pairs[:60] + [(b, a) for a, b in pairs[60:120]]
    pairs = np.array(pairs, dtype=np.int32)
    got = assert_oracle(db, T, p, pairs=pairs)
    assert np.array_equal(got, gpu_scores(db, T, p, pairs=pairs, stack=False))
    assert plan_info(db, T, p, pairs=pairs).stacked_pairs > 0


def test_stacked_strips_in_the_uneven_set_graph_replays():
    """uneven-2000 with stacking and int32 cells: more than half of its pairs
    are scored in stacked strips; graph replays equal the unstacked plan on
    every pair."""
    w = W.uneven_batch(H.tables())
    stacked = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty,
                     opts=N.options(N.KERNEL_AUTO, use_graph=True, packed=False))
    flat = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty,
                  opts=N.options(N.KERNEL_AUTO, stack=False, packed=False))
    out = N.DeviceBuffer(4 * w.npairs)
    try:
        assert stacked.info().stacked_pairs > w.npairs // 2
        flat.execute(out.ptr)
        exp = out.download(np.int32, w.npairs)
        for _ in range(3):
            out.upload(np.full(w.npairs, -7, dtype=np.int32))
            stacked.execute(out.ptr)
            assert np.array_equal(out.download(np.int32, w.npairs), exp)
    finally:
        out.close()
        stacked.close()
        flat.close()


def test_long_pair_beyond_the_line_limit_is_refused():
    """The long kernel addresses each strip-boundary line through one buffer
    descriptor whose byte size must fit 31 bits: a long pair whose target
    needs a longer line is refused at planning (MSA_ERR_TOO_LARGE), before
    anything is allocated, instead of scoring with wrapped offsets.  Two
    sequences of 2^28 residues (one code throughout, blocks of zeros)."""
    T, p = H.tables()["blosum62"]
    nb = (2**28 + 2) // 3
    blocks = np.zeros(2 * nb, dtype=np.uint16)
    offsets = np.array([0, nb], dtype=np.int64)
    sizes = np.array([nb, nb], dtype=np.int32)
    with pytest.raises(N.MsaError) as e:
        N.Plan(blocks, offsets, sizes, T, p, opts=N.options(N.KERNEL_AUTO))
    assert e.value.status == N.MSA_ERR_TOO_LARGE
    assert "boundary line" in str(e.value)


@pytest.mark.parametrize("duo", [True, False])
@pytest.mark.parametrize("table", ["blosum62", "pam250", "default", "asymmetric"])
def test_packed_cells_match_int32_and_oracle(table, duo):
    """Queries that fill one 16 x 32 strip on their own (225..512 rows) are
    scored in packed 16-bit cells, two queries against one target per lane
    group (duo) or one query against two targets: every score equals the
    oracle's and the all-int32 plan's, with odd target counts (a lane group
    with one target) and targets of every length."""
    rng = np.random.default_rng(71 + len(table))
    if table == "asymmetric":
        T, p = H.tables()["blosum62"]
        T = T.copy()
        T[0, 1], T[1, 0], T[6, 13], T[19, 4] = 7, -3, 2, -4
    else:
        T, p = H.tables()[table]
    lengths = list(rng.integers(225, 513, 30)) + list(rng.integers(1, 1500, 13))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    got = assert_oracle(db, T, p, packed=True, duo=duo)
    assert np.array_equal(got, gpu_scores(db, T, p, packed=False))
    info = plan_info(db, T, p, packed=True, duo=duo)
    assert info.packed_pairs > 0
    assert (info.duo_pairs > 0) == duo
    assert plan_info(db, T, p, packed=False).packed_pairs == 0
    # a plan this small keeps its single launch unless packing is asked for
    assert plan_info(db, T, p).packed_pairs == 0


@pytest.mark.parametrize("duo", [True, False])
def test_packed_cells_at_the_top_of_their_range(duo):
    """A packed cell holds at most 0x7BFF: with max(sxx) = 61 a 512-row query
    is packed (61 x 512 = 31,232), and a pair of two identical all-W
    sequences reaches that bound; with max(sxx) = 62 (62 x 512 = 31,744) the
    same query stays on int32 cells.  Scores equal the oracle's either way."""
    T, _ = H.tables()["blosum62"]
    w = W.letter_code("W")
    assert T[w, w] == 11 and T.max() == 11
    rng = np.random.default_rng(5)
    seqs = [np.full(512, w, np.uint8), np.full(512, w, np.uint8), np.full(300, w, np.uint8)]
    seqs += [rng.integers(0, 24, int(n)).astype(np.uint8) for n in rng.integers(225, 513, 6)]
    db = W.pack(seqs)
    for p, packed in ((25, True), (26, False)):
        got = assert_oracle(db, T, p, packed=True, duo=duo)
        assert int(got[0]) == 512 * 11             # pair (1, 0): the identical all-W pair
        info = plan_info(db, T, p, packed=True, duo=duo)
        if packed:
            assert info.packed_pairs > 0
        else:
            # 62 x 512 > 0x7BFF: the 512-row queries stay on int32 cells
            # (shorter ones may still be packed).
            assert np.array_equal(got, gpu_scores(db, T, p, packed=False))
    # 62 x 300 fits: the 300-row query is packed under p = 26 too
    assert plan_info(db, T, 26, packed=True).packed_pairs > 0


def test_packed_cells_in_the_uneven_set():
    """uneven-2000: the default plan scores a share of its pairs in packed
    cells, and every one of the 1,999,000 scores equals the all-int32 plan's
    (graph replays included)."""
    w = W.uneven_batch(H.tables())
    packed = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty, opts=N.options(N.KERNEL_AUTO, use_graph=True))
    flat = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty, opts=N.options(N.KERNEL_AUTO, packed=False))
    out = N.DeviceBuffer(4 * w.npairs)
    try:
        info = packed.info()
        assert info.packed_pairs > 100000
        # nearly every packed query of 225 rows and more finds a partner; the
        # stacks of shorter queries take packed cells too
        assert info.duo_pairs > 0.9 * (info.packed_pairs - info.stacked_pairs)
        assert info.stacked_pairs > 0 and info.packed_pairs > info.duo_pairs + 0.9 * info.stacked_pairs
        assert 32 * 50 * info.packed_pairs <= info.packed_cells <= info.cells   # queries >= 32 rows, targets >= 50
        assert flat.info().packed_pairs == 0 and flat.info().packed_cells == 0
        flat.execute(out.ptr)
        exp = out.download(np.int32, w.npairs)
        for _ in range(2):
            out.upload(np.full(w.npairs, -7, dtype=np.int32))
            packed.execute(out.ptr)
            assert np.array_equal(out.download(np.int32, w.npairs), exp)
    finally:
        out.close()
        packed.close()
        flat.close()


@pytest.mark.parametrize("duo", [True, False])
def test_packed_cells_longer_than_a_strip_at_the_top_of_their_range(duo):
    """Queries longer than one strip take packed cells when their rebased
    strips' cells fit: max(sxx) x 592 <= 0x7BFF.  All-W sequences of 2000 and
    3000 residues with max(sxx) = 53 (blosum62, gap 21) make every cell grow
    as fast as it can, so each strip's base is raised many times close to the
    bound; with max(sxx) = 54 they stay on int32 cells.  Scores equal the
    oracle's and the int32 plan's either way."""
    T, _ = H.tables()["blosum62"]
    w = W.letter_code("W")
    rng = np.random.default_rng(9)
    seqs = [np.full(2000, w, np.uint8), np.full(3000, w, np.uint8), np.full(3000, w, np.uint8)]
    seqs += [rng.integers(0, 24, int(n)).astype(np.uint8) for n in (700, 1300, 2600)]
    db = W.pack(seqs)
    for p, packed in ((21, True), (22, False)):
        got = assert_oracle(db, T, p, packed=True, duo=duo)
        assert int(got[2]) == 3000 * 11           # pair (2, 1): identical all-W 3000-residue sequences
        assert np.array_equal(got, gpu_scores(db, T, p, packed=False))
        info = plan_info(db, T, p, packed=True, duo=duo)
        assert (info.packed_pairs > 0) == packed
        if packed:                                # every query of the three all-W pairs is longer than 2000
            assert info.packed_cells >= 2000 * 3000 * 2 + 3000 * 3000


@pytest.mark.parametrize("table", ["blosum62", "asymmetric"])
def test_duo_items_with_explicit_pair_lists(table):
    """Duo items from an explicit pair list: queries whose target lists only
    partly agree (entries -1 for the half that does not pair), a pair named
    twice, both orientations of an asymmetric table in one duo, queries of
    different strip counts sharing items (the shorter padded with neutral
    rows on top), and queries left alone (high halves idle).  Every score
    equals the oracle's and the int32 plan's."""
    rng = np.random.default_rng(23)
    T, p = H.tables()["blosum62"]
    if table == "asymmetric":
        T = T.copy()
        T[0, 1], T[1, 0], T[6, 13], T[19, 4] = 7, -3, 2, -4
    lengths = [300, 310, 505, 530, 1030, 1100, 2100, 2150, 700, 260, 990, 1500]
    lengths += list(rng.integers(600, 2400, 28))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    n = len(lengths)
    pairs = [(i, j) for i in range(n) for j in range(i) if rng.random() < 0.8]
    pairs += [(5, 4), (4, 5), (7, 6), (3, 2), (3, 2)]
    pairs = np.array(pairs, dtype=np.int32)
    got = assert_oracle(db, T, p, pairs=pairs, packed=True)
    assert np.array_equal(got, gpu_scores(db, T, p, pairs=pairs, packed=False))
    assert np.array_equal(got, gpu_scores(db, T, p, pairs=pairs, packed=True, duo=False))
    info = plan_info(db, T, p, pairs=pairs, packed=True)
    assert info.duo_pairs > 0


def test_duo_items_in_graph_replays_after_other_plans():
    """A duo plan replayed as a graph after plans of the other kernels ran
    on the same stream scores the same every time (its lines and queue
    counters start clean)."""
    T, p = H.tables()["blosum62"]
    rng = np.random.default_rng(31)
    lengths = list(rng.integers(230, 1800, 60))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    exp = H.oracle_scores(db, T, p)
    duo = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=True, packed=True))
    other = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, packed=True, duo=False))
    out = N.DeviceBuffer(4 * len(exp))
    try:
        assert duo.info().duo_pairs > 0
        for _ in range(3):
            out.upload(np.full(len(exp), -7, dtype=np.int32))
            other.execute(out.ptr)
            assert np.array_equal(out.download(np.int32, len(exp)).astype(np.float32), exp)
            out.upload(np.full(len(exp), -7, dtype=np.int32))
            duo.execute(out.ptr)
            assert np.array_equal(out.download(np.int32, len(exp)).astype(np.float32), exp)
    finally:
        out.close()
        duo.close()
        other.close()


def test_duo_launch_with_multi_strip_int32_items():
    """With max(sxx) = 55 (blosum62, gap 22) one-strip queries fit 16-bit
    cells and take duo items, while queries longer than a strip do not
    (55 x 592 > 0x7BFF) and take int32 items: the duo launch scores those
    after its duo items, each wave in its own part of the workgroup's LDS and
    with its own strip-boundary lines.  Every score equals the oracle's and
    the all-int32 plan's, over graph replays too."""
    T, _ = H.tables()["blosum62"]
    p = 22
    rng = np.random.default_rng(41)
    lengths = list(rng.integers(230, 512, 36)) + list(rng.integers(600, 1700, 12))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    info = plan_info(db, T, p, packed=True)
    assert info.duo_pairs > 0
    assert info.batched_pairs > info.packed_pairs + info.stacked_pairs   # int32 items beside the duo items
    got = assert_oracle(db, T, p, packed=True)
    assert np.array_equal(got, gpu_scores(db, T, p, packed=False))
    plan = N.Plan(*db, T, p, opts=N.options(N.KERNEL_AUTO, use_graph=True, packed=True))
    out = N.DeviceBuffer(4 * len(got))
    try:
        for _ in range(3):
            out.upload(np.full(len(got), -7, dtype=np.int32))
            plan.execute(out.ptr)
            assert np.array_equal(out.download(np.int32, len(got)), got)
    finally:
        out.close()
        plan.close()


@pytest.mark.parametrize("table", ["blosum62", "asymmetric"])
def test_packed_stacked_strips(table):
    """With duo items in the plan, stacks of short queries take packed cells
    (two targets per lane group): members in both orientations of an
    asymmetric table, members whose target lists differ (explicit pairs),
    pairs named twice, odd target counts.  Every score equals the oracle's
    and the all-int32 plan's."""
    rng = np.random.default_rng(57)
    T, p = H.tables()["blosum62"]
    if table == "asymmetric":
        T = T.copy()
        T[0, 1], T[1, 0], T[6, 13], T[19, 4] = 7, -3, 2, -4
    lengths = list(rng.integers(32, 220, 40)) + list(rng.integers(240, 500, 24)) + list(rng.integers(520, 900, 6))
    db = W.pack([rng.integers(0, 24, int(n)).astype(np.uint8) for n in lengths])
    n = len(lengths)
    pairs = [(i, j) for i in range(n) for j in range(i) if rng.random() < 0.85]
    pairs += [(3, 1), (3, 1), (50, 7)]
    pairs = np.array(pairs, dtype=np.int32)
    info = plan_info(db, T, p, pairs=pairs, packed=True)
    assert info.duo_pairs > 0 and info.stacked_pairs > 0
    assert info.packed_pairs >= info.duo_pairs + info.stacked_pairs   # every stack in packed cells
    got = assert_oracle(db, T, p, pairs=pairs, packed=True)
    assert np.array_equal(got, gpu_scores(db, T, p, pairs=pairs, packed=False))
