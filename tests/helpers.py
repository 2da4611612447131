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


# ---- FASTA texts --------------------------------------------------------------

# Edge cases of src/io/loader/parser/fasta.cpp:26-70 (getline / peek rules).
FASTA_EDGE_CASES = [
    b"", b">", b">\n", b"\n", b"\n\n", b"ACGT", b"ACGT\n", b">a\nACGT", b">a\nACGT\n", b">a\nACGT\n\n",
    b">a\nAC\n\nGT\n>b\nTT", b"junk\n>a\nA\n>b\n\n>c\nAAAA\r\nCC\r\n", b">a\n>b\n>c", b">a\n>b\n>c\n",
    b"\n\n>x\nA\nB\nC\n\nD\n>y\n-", b">only header, no newline", b">h\r\nAC\r\n\r\nGT\r\n", b">h\nA>B\nC",
    b">h\nABC\nDEF\nG", b">h\nAB\nCD\nEF\n", b">h\nA\nB\nC\nD\nE\n", b">\n\n\n>\nX", b"\x00>\nA\x00C\xff",
    b">a\nacgtn*xyz-09\n", b" >not a header\n>yes\nA\n", b">a\n\n\n\nTT\n>b\nGG",
]


def random_fasta_bytes(rng, length, alphabet=b"ACGTacgtNn-*>\n\r xyz09\x00\xff>"):
    """Random bytes over an alphabet rich in the parser's special bytes."""
    if length <= 0:
        return b""
    return bytes(rng.choice(np.frombuffer(alphabet, np.uint8), int(length)).astype(np.uint8))


def synthetic_fasta(rng, records, lo, hi, width=60, crlf=False, lower=False, blank_every=0, junk=False,
                    letters=b"ARNDCQEGHILKMFPSTWYVBZX"):
    """A FASTA text of `records` records with lengths U[lo, hi], lines of
    `width` residues (0: one line per record), optional CRLF line ends,
    lower-case residues, an empty line after every `blank_every`-th record and
    junk lines before the first header and after empty lines."""
    nl = b"\r\n" if crlf else b"\n"
    alpha = np.frombuffer(letters.lower() if lower else letters, np.uint8)
    parts = []
    if junk:
        parts.append(b"preamble line" + nl)
    for r in range(records):
        parts.append(b">seq%d description %d" % (r, int(rng.integers(0, 1 << 30))) + nl)
        n = int(rng.integers(lo, hi + 1))
        seq = bytes(rng.choice(alpha, n).astype(np.uint8)) if n else b""
        w = width if width > 0 else max(n, 1)
        for k in range(0, n, w):
            parts.append(seq[k:k + w] + nl)
        if blank_every and (r + 1) % blank_every == 0:
            parts.append(nl if not crlf else b"\n")
            if junk:
                parts.append(b"junk after an empty line" + nl)
    return b"".join(parts)


def decode_store(blocks, offset, size):
    """Decoded (padded) contents of every record of a sequence store."""
    return [M.decode([int(b) for b in blocks[o:o + s]]) for o, s in zip(offset, size)]
