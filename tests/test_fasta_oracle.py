"""The FASTA oracle (oracle/fasta_oracle.c), pinned on the CPU.

Pins: the reference's own parser compiled from its sources
(oracle/_ref/libref_fasta.so: src/io/loader/parser/fasta.cpp, encoder.cpp,
database.cpp), the md5 digests of test/loader.py, and the Python restatement
(parse_fasta + encode) on the reference's fixtures.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import museqa_ref as M
from tests import helpers as H

ORACLE = M.FastaOracle()
needs_ref = pytest.mark.skipif(not M.ReferenceFasta.available(), reason="oracle/_ref not built here")


def assert_same_as_reference(text, tmp_path):
    path = tmp_path / "case.fasta"
    path.write_bytes(text)
    ref = M.ReferenceFasta().load(str(path))
    blocks, offset, size, ds, de = ORACLE.parse(text)
    assert ref is not None
    rb, ro, rs, rd = ref
    assert len(ro) == len(offset), text[:80]
    assert np.array_equal(rb, blocks) and np.array_equal(ro, offset) and np.array_equal(rs, size), text[:80]
    assert rd == [text[a:b] for a, b in zip(ds, de)]


@needs_ref
@pytest.mark.parametrize("k", range(len(H.FASTA_EDGE_CASES)))
def test_edge_cases_match_the_reference_build(k, tmp_path):
    assert_same_as_reference(H.FASTA_EDGE_CASES[k], tmp_path)


@needs_ref
def test_random_texts_match_the_reference_build(tmp_path):
    rng = np.random.default_rng(7)
    for _ in range(600):
        assert_same_as_reference(H.random_fasta_bytes(rng, rng.integers(0, 300)), tmp_path)


@needs_ref
@pytest.mark.parametrize("crlf,lower,blank,width", [(False, False, 0, 60), (True, True, 3, 7), (False, False, 2, 0)])
def test_synthetic_files_match_the_reference_build(crlf, lower, blank, width, tmp_path):
    rng = np.random.default_rng(11)
    text = H.synthetic_fasta(rng, 40, 0, 700, width=width, crlf=crlf, lower=lower, blank_every=blank, junk=True)
    assert_same_as_reference(text, tmp_path)


def test_fixtures_match_the_python_restatement_and_the_loader_digests():
    with open(os.path.join(H.GOLDEN, "loader_md5.json")) as f:
        digests = json.load(f)["digests"]
    for fixture, expected in digests.items():
        with open(os.path.join(H.GOLDEN, fixture), "rb") as f:
            text = f.read()
        blocks, offset, size, ds, de = ORACLE.parse(text)
        recs = M.parse_fasta(os.path.join(H.GOLDEN, fixture))
        assert [text[a:b].decode() for a, b in zip(ds, de)] == [d for d, _ in recs]
        decoded = H.decode_store(blocks, offset, size)
        assert decoded == [M.decode(M.encode(s)) for _, s in recs]
        by_key = dict(zip([d for d, _ in recs], decoded))
        for key, digest in expected.items():
            assert hashlib.md5(by_key[key].encode()).hexdigest() == digest, (fixture, key)


def test_structure_of_a_synthetic_file():
    """Record count, sizes ceil(len/3) and contents of a file built here."""
    rng = np.random.default_rng(5)
    lengths = [0, 1, 2, 3, 4, 59, 60, 61, 179, 180, 181]
    seqs = [bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), n).astype(np.uint8)) for n in lengths]
    text = b"".join(b">r%d\n" % i + b"".join(s[k:k + 60] + b"\n" for k in range(0, len(s), 60))
                    for i, s in enumerate(seqs))
    blocks, offset, size, ds, de = ORACLE.parse(text)
    assert list(size) == [(n + 2) // 3 for n in lengths]
    assert list(offset) == list(np.concatenate([[0], np.cumsum(size)[:-1]]))
    for s, got in zip(seqs, H.decode_store(blocks, offset, size)):
        assert got[:len(s)] == s.decode() and set(got[len(s):]) <= {"*"}


@pytest.mark.parametrize("k", range(1, 7))
def test_reference_benchmark_databases(k, tmp_path):
    """db/case<k>.fasta (kept compressed under tests/golden/db/, sha256 of the
    original checked on load): the oracle's store has one sequence per '>'
    record, each of the residues its lines hold, and equals the reference's
    own parser where that is built."""
    from msa_amd import workloads as W
    text = W.reference_case_text(k)
    records = []
    for line in text.split(b"\n"):
        if line.startswith(b">"):
            records.append(0)
        elif records:
            records[-1] += len(line.strip())
    blocks, offset, size, _, _ = ORACLE.parse(text)
    assert len(size) == len(records)
    assert np.array_equal(size, [(r + 2) // 3 for r in records])
    if M.ReferenceFasta.available():
        assert_same_as_reference(text, tmp_path)
