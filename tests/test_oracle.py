"""Pins the oracle (CPU, no GPU needed).

The oracle is trusted only after it reproduces, independently:
  - the md5 digests the reference's loader test holds for its FASTA fixtures
    (test/loader.py:14-36) -> FASTA parser + encoder + decoder;
  - the reference's own sequential.cpp / encoder.cpp, compiled from the
    reference tree into oracle/_ref (where that tree exists);
  - textbook global alignment (full matrix) on clean inputs, closed forms and
    invariants;
  - the committed golden scores of the reference's pairwise fixtures.
"""
import hashlib
import json
import os
import re

import numpy as np
import pytest

from oracle import museqa_ref as M
from msa_amd import workloads as W
from tests import helpers as H

REF_TABLE_SRC = "/root/reference/src/pairwise/table.cu"


def test_loader_md5_digests():
    """test/loader.py:39-52: md5(str(contents)) of every record."""
    with open(os.path.join(H.GOLDEN, "loader_md5.json")) as f:
        digests = json.load(f)["digests"]
    for fixture, expected in digests.items():
        recs = dict(M.parse_fasta(os.path.join(H.GOLDEN, fixture)))
        for key, digest in expected.items():
            decoded = M.decode(M.encode(recs[key]))
            assert hashlib.md5(decoded.encode()).hexdigest() == digest, (fixture, key)


SEQUENCE_FIXTURES = [   # test/sequence.py:11-17
    "MNNQRKKTGRPSFNMLKRARNRVSTGSQLAKRFSKGLLSGQGPMKLVMAFIAFLRFLAIPPTAGILARWS",
    "TEVKGYTKGIDJAOIJDOADUIHDAIOAFAOIFJAOIFJOAIJFOAICJOAIMCOQIOIDUAITGYRQQ",
    "MKNPKKKSGGFRIVNMLKRGVARVSPFGGLKRLPAGLLLGHGPIRMVLAILAFLRFTAIKPSLGLINRW",
    "AAAAATCGATCAAGCAATCTCTCAGTCTCTCAGTTTACAGCAATCTTTGCAAACACACGAACCACAAA",
    "LLNRFTMAHMKPTYERDVDLGAGTRHVAVEPEVANLDIIGQRIENIKNEHKSTWHYDEDNPYKTWAY",
]

# Letter -> code as the reference's encoder documents it (src/encoder.cpp:27-39).
LETTER_CODES = dict(A=0, B=20, C=1, D=6, E=8, F=14, G=3, H=9, I=10, J=21, K=12, L=11, M=13, N=5, O=23,
                    P=15, Q=7, R=4, S=16, T=2, U=23, V=19, W=17, X=23, Y=18, Z=22)


@pytest.mark.parametrize("text", SEQUENCE_FIXTURES)
def test_sequence_round_trip(text):
    """test/sequence.py:26-48: length = padded length, per-index decode, str()."""
    blocks = M.encode(text)
    padded = 3 * len(blocks)
    assert padded == len(text) + (-len(text)) % 3
    decoded = M.decode(blocks)
    expected = "".join("ACTGRNDQEHILKMFPSWYVBJZX*-"[LETTER_CODES[c]] for c in text) + "*" * (padded - len(text))
    assert decoded == expected
    for i, ch in enumerate(text):
        assert M.access(blocks, i) == LETTER_CODES[ch]


def test_encoder_block_layout():
    """Closed form of the block layout: codes at shifts 1/6/11, bit 0 marks padding."""
    assert M.encode("") == []
    assert M.encode("A") == [(0 << 1) | (24 << 6) | (24 << 11) | 1]
    assert M.encode("ACT") == [(0 << 1) | (1 << 6) | (2 << 11)]
    assert M.encode("ACTG") == [(0 << 1) | (1 << 6) | (2 << 11), (3 << 1) | (24 << 6) | (24 << 11) | 1]
    assert M.encode_char("-") == 25 and M.encode_char("*") == 24 and M.encode_char("1") == 24
    assert M.encode_char("a") == 0 and M.encode_char("\r") == 24


@pytest.mark.skipif(not M.ReferenceBuild.available(), reason="reference build (oracle/_ref) not present")
def test_encoder_matches_reference_build():
    ref = M.ReferenceBuild()
    rng = np.random.default_rng(3)
    pool = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefxyz*-.1 \r"
    for _ in range(200):
        text = "".join(pool[i] for i in rng.integers(0, len(pool), int(rng.integers(0, 40))))
        assert ref.encode(text) == M.encode(text), repr(text)


def test_fasta_parser_rules(tmp_path):
    """fasta.cpp:26-45: records end at '>' or at the first empty line; text
    before the first '>' is skipped; a CR stays in the line."""
    p = tmp_path / "x.fasta"
    p.write_bytes(b"junk\n>a\nAC\nGT\n\nTT\n>b desc\r\nAA\r\n>c\n")
    assert M.parse_fasta(str(p)) == [("a", "ACGT"), ("b desc\r", "AA\r"), ("c", "")]


def test_golden_scores_reproduce():
    for key, entry in H.golden_scores().items():
        fixture, table = key.split("/")
        _, db = H.fixture_db(fixture)
        T, p = H.tables()[table]
        got = H.oracle_scores(db, T, p)
        assert [int(v) for v in got] == entry["scores"], key


def test_oracle_matches_textbook_on_clean_input():
    rng = np.random.default_rng(1)
    tables = H.tables()
    for trial in range(60):
        name = list(tables)[trial % len(tables)]
        T, p = tables[name]
        strings = ["".join("ACDEFGHIKLMNPQRSTVWYBZXJ"[c] for c in rng.integers(0, 24, int(rng.integers(0, 50))))
                   for _ in range(4)]
        db = H.strings_db(strings)
        got = H.oracle_scores(db, T, p, workers=1)
        for c, (i, j) in enumerate(M.pair_order(4)):
            one, two = (i, j) if db[2][i] > db[2][j] else (j, i)
            a = [M.encode_char(ch) for ch in strings[one]]
            b = [M.encode_char(ch) for ch in strings[two]]
            assert got[c] == M.textbook_nw(a, b, T, p), (name, strings[i], strings[j])


def test_oracle_matches_python_restatement_on_padding_codes():
    """Internal padding codes: the row sequence stops at its first one, a
    column holding one copies its left cell (sequential.cpp:40-66)."""
    rng = np.random.default_rng(2)
    tables = H.tables()
    ref = M.ReferenceBuild() if M.ReferenceBuild.available() else None
    for trial in range(40):
        T, p = tables[["blosum62", "pam250", "default"][trial % 3]]
        pool = "ACGTWYK**1-"
        strings = ["".join(pool[i] for i in rng.integers(0, len(pool), int(rng.integers(0, 30)))).replace("-", "A")
                   for _ in range(4)]
        seqs = [M.encode(s) for s in strings]
        db = H.pack_blocks(seqs)
        got = H.oracle_scores(db, T, p, workers=1)
        assert list(got) == [np.float32(v) for v in M.needleman(seqs, T.tolist(), p)]
        if ref is not None:
            pairs = np.array(M.pair_order(4), dtype=np.int32)
            assert np.array_equal(got, ref.align_list(db[0], db[1], db[2], pairs, T, p))


@pytest.mark.skipif(not M.ReferenceBuild.available(), reason="reference build (oracle/_ref) not present")
def test_oracle_orientation_matches_reference_align():
    """The orientation rule -- rows = the sequence with more blocks, on a tie
    pair.second (sequential.cpp:87-91) -- pinned to the reference's own
    align() (oracle/_ref, the TU's file-local align over a museqa::database),
    not to a restatement: 4 databases x 4 asymmetric tables x 66 pairs with
    many equal block counts, internal padding codes and explicit lists that
    hold both orders of a pair.  The same lists with every pair reversed must
    give different scores somewhere, so the check would catch a flipped tie."""
    ref = M.ReferenceBuild()
    rng = np.random.default_rng(41)
    T, p = H.tables()["blosum62"]
    cases = sensitive = 0
    for trial in range(4):
        # Lengths 3k..3k+2 for a few k: many pairs share a block count.
        seqs = []
        for _ in range(12):
            k = int(rng.integers(1, 5)) * 4
            c = rng.integers(0, 24, 3 * k + int(rng.integers(0, 3))).astype(np.uint8)
            if rng.random() < 0.4:
                c[int(rng.integers(0, len(c)))] = 24
            seqs.append(c)
        db = W.pack(seqs)
        pairs = np.array([(i, j) if rng.random() < 0.5 else (j, i) for i in range(12) for j in range(i)],
                         dtype=np.int32)
        for _ in range(4):
            Ta = T.copy()
            for _ in range(12):
                a, b = rng.integers(0, 24, 2)
                Ta[a, b] = int(rng.integers(-4, 12))
            got = H.oracle_scores(db, Ta, p, pairs=pairs, workers=1)
            exp = ref.align_list(db[0], db[1], db[2], pairs, Ta, p, workers=3)
            assert np.array_equal(got, exp)
            rev = H.oracle_scores(db, Ta, p, pairs=pairs[:, ::-1].copy(), workers=1)
            sensitive += int(np.sum(rev != got))
            cases += len(pairs)
    assert cases >= 200
    assert sensitive > 0


def test_closed_forms():
    tables = H.tables()
    T, p = tables["default"]
    for text in ["A", "ACG", "ACGT", "MKVLAAGW" * 5]:
        db = H.strings_db([text, text])
        assert H.oracle_scores(db, T, p)[0] == len(text)           # identity under +1/-1
        db = H.strings_db([text, ""])
        assert H.oracle_scores(db, T, p)[0] == -p * len(text)       # versus empty
    T, p = tables["blosum62"]
    db = H.strings_db(["", ""])
    assert H.oracle_scores(db, T, p)[0] == 0
    # A row sequence that starts with a padding code runs no row: row 0 at the
    # padded width of the column sequence is the result.
    db = H.strings_db(["*AAAAAA", "ACGT"])
    assert H.oracle_scores(db, T, p)[0] == -p * 6


def test_planted_mistakes_are_caught():
    """A wrong gap sign or a transposed lookup in a copy of the recurrence
    fails the textbook check, so that check has teeth."""
    T, p = H.tables()["pam250"]
    a = [M.encode_char(c) for c in "WWGAEK"]
    b = [M.encode_char(c) for c in "GAWKE"]
    good = M.textbook_nw(a, b, T, p)
    assert M.textbook_nw(a, b, T, -p) != good
    Tt = T.copy()
    Tt[17, 3], Tt[3, 17] = 9, -9
    assert M.textbook_nw(a, b, Tt, p) != M.textbook_nw(b, a, Tt, p)


def test_pair_order_closed_form():
    for n in range(0, 9):
        order = M.pair_order(n)
        assert len(order) == n * (n - 1) // 2
        for c, (i, j) in enumerate(order):
            assert c == i * (i - 1) // 2 + j and j < i


def test_tables_hash_and_quirks():
    with open(os.path.join(os.path.dirname(M.__file__), "tables.json")) as f:
        data = json.load(f)
    h = hashlib.sha256()
    for t in data["tables"]:
        h.update(t["name"].encode())
        h.update(t["penalty"].to_bytes(1, "little", signed=True))
        h.update(bytes((v & 0xFF) for row in t["matrix"] for v in row))
    assert h.hexdigest() == data["sha256"] == "27ed98dfee7a79c2ac6be712bd1d91727cd75e52c50674f4be320663f0e1e59b"
    tables = M.load_tables()
    assert [n for n in tables] == ["default", "blosum62", "blosum45", "blosum50", "blosum80", "blosum90", "pam250"]
    assert [tables[n][1] for n in tables] == [1, 4, 5, 5, 6, 6, 8]
    b62 = tables["blosum62"][0]
    A, C, W_, G, E, X = 0, 1, 17, 3, 8, 23
    assert b62[A, A] == 4 and b62[C, C] == 9 and b62[W_, W_] == 11 and b62[X, A] == -1
    pam = tables["pam250"][0]
    assert pam[G, G] == 4 and pam[E, E] == 5          # G and E exchanged w.r.t. textbook PAM250
    assert all((m == m.T).all() for m, _ in tables.values())


@pytest.mark.skipif(not os.path.exists(REF_TABLE_SRC), reason="reference tree not present")
def test_tables_match_reference_source():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gen_tables", os.path.join(H.GOLDEN, "..", "..", "msa_amd", "csrc", "gen_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(REF_TABLE_SRC) as f:
        entries = gen.parse_reference_tables(f.read())
    tables = M.load_tables()
    for name, penalty, matrix in entries:
        assert tables[name][1] == penalty
        assert (tables[name][0] == np.array(matrix)).all()
    with open(os.path.join(os.path.dirname(M.__file__), "..", "msa_amd", "csrc", "scoring_tables.inc")) as f:
        inc = f.read()
    assert gen.digest(entries) in inc
    assert re.search(r"kTableCount = 7", inc)
