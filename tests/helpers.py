"""Shared helpers for the tests: fixture loading, random databases, oracle calls."""
import json
import os

import numpy as np

from oracle import museqa_ref as M
from msa_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tables():
    return {k: (v[0].astype(np.int32), v[1]) for k, v in M.load_tables().items()}


def fixture_db(name):
    recs = M.parse_fasta(os.path.join(GOLDEN, name))
    seqs = [M.encode(s) for _, s in recs]
    return recs, pack_blocks(seqs)


def pack_blocks(seqs):
    """Block lists -> (blocks, offsets, sizes) arrays."""
    blocks = np.array([b for s in seqs for b in s], dtype=np.uint16)
    sizes = np.array([len(s) for s in seqs], dtype=np.int32)
    offsets = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs):
        offsets[1:] = np.cumsum(sizes, dtype=np.int64)[:-1]
    return blocks, offsets, sizes


def golden_scores():
    with open(os.path.join(GOLDEN, "mock_scores.json")) as f:
        return json.load(f)["scores"]


def oracle_scores(db, table, penalty, pairs=None, workers=8):
    blocks, offsets, sizes = db
    return M.CpuOracle().align(blocks, offsets, sizes, np.asarray(table, dtype=np.float32), penalty,
                               pairs=pairs, workers=workers)


def random_db(rng, n, lo, hi, dirty=False, alphabet=24):
    """n random sequences (codes 0..alphabet-1) with lengths in [lo, hi];
    dirty=True plants padding codes (24) at random positions in half of them."""
    codes = []
    for _ in range(n):
        length = int(rng.integers(lo, hi + 1))
        c = rng.integers(0, alphabet, length).astype(np.uint8)
        if dirty and length and rng.random() < 0.5:
            for _ in range(int(rng.integers(1, 4))):
                c[int(rng.integers(0, length))] = 24
        codes.append(c)
    return W.pack(codes)


def strings_db(strings):
    return pack_blocks([M.encode(s) for s in strings])


def assert_same_tree(got, exp):
    assert len(got) == len(exp)
    for field in ("parent", "child", "level"):
        bad = np.nonzero(np.any((got[field] != exp[field]).reshape(len(got), -1), axis=1))[0]
        assert len(bad) == 0, "%s differs at nodes %s: gpu %s oracle %s" % (
            field, bad[:5], got[field][bad[:5]], exp[field][bad[:5]])
    g, e = got["distance"], exp["distance"]
    same = (g.view(np.uint32) == e.view(np.uint32)) | (np.isnan(g) & np.isnan(e))
    bad = np.nonzero(~same)[0]
    assert len(bad) == 0, "distance differs at nodes %s: gpu %s oracle %s" % (bad[:5], g[bad[:5]], e[bad[:5]])
