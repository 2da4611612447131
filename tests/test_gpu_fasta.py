"""FASTA loading on the GPU (msa_fasta_*) against the oracle
(oracle/fasta_oracle.c, itself pinned to the reference's own parser and the
loader digests of test/loader.py).  Every store must be identical: blocks,
offsets, sizes and description ranges."""
import hashlib
import json
import os

import numpy as np
import pytest

from msa_amd import _native as N
from oracle import museqa_ref as M
from tests import helpers as H

pytestmark = pytest.mark.gpu

ORACLE = M.FastaOracle()


@pytest.fixture(scope="module")
def loader():
    ld = N.FastaLoader()
    yield ld
    ld.close()


def assert_store_equal(loader, text, label=""):
    info = loader.parse(text)
    got = loader.store()
    exp = ORACLE.parse(text)
    assert info["sequences"] == len(exp[1]), (label, info, len(exp[1]))
    assert info["blocks"] == len(exp[0]), (label, info)
    for name, g, e in zip(("blocks", "offset", "size", "desc_start", "desc_end"), got, exp):
        if not np.array_equal(g, e):
            bad = np.nonzero(g != e)[0] if len(g) == len(e) else []
            raise AssertionError("%s %s differs (%d vs %d entries) first at %s" % (label, name, len(g), len(e),
                                                                                  bad[:5]))
    return info


def test_fixtures_and_loader_digests(loader):
    with open(os.path.join(H.GOLDEN, "loader_md5.json")) as f:
        digests = json.load(f)["digests"]
    for fixture, expected in digests.items():
        with open(os.path.join(H.GOLDEN, fixture), "rb") as f:
            text = f.read()
        assert_store_equal(loader, text, fixture)
        blocks, offset, size, ds, de = loader.store()
        names = [text[a:b].decode() for a, b in zip(ds, de)]
        decoded = dict(zip(names, H.decode_store(blocks, offset, size)))
        for key, digest in expected.items():
            assert hashlib.md5(decoded[key].encode()).hexdigest() == digest, (fixture, key)


@pytest.mark.parametrize("k", range(len(H.FASTA_EDGE_CASES)))
def test_edge_cases(loader, k):
    assert_store_equal(loader, H.FASTA_EDGE_CASES[k], repr(H.FASTA_EDGE_CASES[k][:40]))


def test_random_small_texts(loader):
    rng = np.random.default_rng(3)
    for i in range(800):
        text = H.random_fasta_bytes(rng, rng.integers(0, 400))
        assert_store_equal(loader, text, "random small %d" % i)


@pytest.mark.parametrize("seed,length", [(1, 17000), (2, 50000), (3, 200000)])
def test_random_texts_across_tiles(loader, seed, length):
    """Random bytes with short lines: many line starts per 64-byte chunk and
    blocks that straddle chunk and tile boundaries."""
    rng = np.random.default_rng(seed)
    assert_store_equal(loader, H.random_fasta_bytes(rng, length), "seed %d" % seed)


@pytest.mark.parametrize("records,lo,hi,width,crlf,lower,blank,junk", [
    (500, 0, 3000, 60, False, False, 0, False),
    (300, 100, 2000, 80, True, False, 7, True),
    (2000, 0, 5, 60, False, True, 0, False),          # tiny records, many per tile
    (40, 20000, 40000, 0, False, False, 0, False),     # one line per record, lines across many tiles
    (200, 1, 900, 1, False, False, 5, True),           # one residue per line
    (100, 0, 4000, 61, True, True, 3, True),
])
def test_synthetic_files(loader, records, lo, hi, width, crlf, lower, blank, junk):
    rng = np.random.default_rng(records + lo + hi + width)
    text = H.synthetic_fasta(rng, records, lo, hi, width=width, crlf=crlf, lower=lower, blank_every=blank,
                             junk=junk)
    info = assert_store_equal(loader, text, "synthetic")
    assert info["bytes"] == len(text)


def test_text_already_on_the_device_at_unaligned_addresses(loader):
    rng = np.random.default_rng(9)
    text = H.synthetic_fasta(rng, 200, 0, 1500, width=70, crlf=True, blank_every=11, junk=True)
    buf = N.DeviceBuffer(len(text) + 64)
    exp = ORACLE.parse(text)
    for shift in (0, 1, 3, 8, 13):
        buf.upload(np.frombuffer(b"\0" * shift + text, dtype=np.uint8))
        loader.parse_device(buf.ptr + shift, len(text))
        for g, e in zip(loader.store(), exp):
            assert np.array_equal(g, e), shift
    buf.close()


def test_buffers_are_reused_across_sizes(loader):
    rng = np.random.default_rng(4)
    big = H.synthetic_fasta(rng, 800, 0, 2500, width=60)
    small = H.synthetic_fasta(rng, 3, 1, 50, width=60)
    for text in (big, small, b"", big[: len(big) // 3], small):
        assert_store_equal(loader, text, "reuse %d" % len(text))


def test_store_feeds_the_pairwise_stage(loader):
    """Parse on the GPU, align the store: scores equal those of the host-packed
    database (the fixtures' golden scores)."""
    golden = H.golden_scores()
    for fixture in ("mock1.fasta", "mock2.fasta", "mock3.fasta"):
        with open(os.path.join(H.GOLDEN, fixture), "rb") as f:
            loader.parse(f.read())
        blocks, offset, size, _, _ = loader.store()
        t, p = N.table("blosum62")
        scores = N.align(blocks, offset, size, t, p)
        assert [int(v) for v in scores] == golden["%s/blosum62" % fixture]["scores"], fixture


def _entries(db):
    return [(e.description, str(e.contents)) for e in (db[i] for i in range(db.count))]


def test_database_load_on_device_matches_the_host_loader(tmp_path):
    """museqa::database::load_on_device builds the database the host loader
    (the reference's rules) builds: same records, descriptions and contents,
    same key rule for repeated descriptions."""
    from msa_amd.museqa.database import Database
    for fixture in ("mock1.fasta", "mock2.fasta", "mock3.fasta"):
        path = os.path.join(H.GOLDEN, fixture)
        assert _entries(Database.load(path, device=True)) == _entries(Database.load(path))
    with open(os.path.join(H.GOLDEN, "loader_md5.json")) as f:
        digests = json.load(f)["digests"]
    for fixture, expected in digests.items():
        db = Database.load(os.path.join(H.GOLDEN, fixture), device=True)
        for key, digest in expected.items():
            assert digest == hashlib.md5(str(db[key].contents).encode()).hexdigest()
    rng = np.random.default_rng(21)
    text = H.synthetic_fasta(rng, 300, 0, 2500, width=70, crlf=True, blank_every=9, junk=True)
    text += b">seq5 repeated description\nACGT\n"
    path = tmp_path / "synthetic.fasta"
    path.write_bytes(text)
    host, dev = Database.load(str(path)), Database.load(str(path), device=True)
    assert host.count == dev.count and _entries(host) == _entries(dev)


def test_pipeline_from_fasta_text_to_guide_tree(tmp_path):
    """The reference's pipeline with every stage on the GPU: the FASTA file
    parsed on the device (database::load_on_device), all pairs scored by the
    pairwise module (src/pairwise.cpp:27-55), the tree built by the phylogeny
    module (src/phylogeny.cpp:25-47).  Against the oracles end to end: the
    oracle's store of the same text, the oracle's scores of that store, the
    oracle's neighbor joining of those scores."""
    from msa_amd.museqa import phylogeny as mpg
    from msa_amd.museqa._core import core
    from msa_amd.museqa.database import Database
    rng = np.random.default_rng(29)
    text = H.synthetic_fasta(rng, 48, 20, 600, width=60, crlf=False, blank_every=7, junk=True)
    path = tmp_path / "pipeline.fasta"
    path.write_bytes(text)
    db = Database.load(str(path), device=True)
    out_db, dm, count = core.module_run({"scoring-table": "blosum62"}, db.native)
    blocks, offset, size, _, _ = ORACLE.parse(text)
    assert count == len(size) == 48
    T, p = H.tables()["blosum62"]
    scores = H.oracle_scores((blocks, offset, size), T, p)
    got_scores = np.array([dm.at(i, j) for i in range(count) for j in range(i)], dtype=np.float32)
    assert np.array_equal(got_scores, scores)
    tdb, tree = core.phylogeny_module_run({"phylogeny": "default"}, out_db, dm, count)
    H.assert_same_tree(mpg.GuideTree(tree).nodes(), M.NjOracle().build(scores, count, 0))
