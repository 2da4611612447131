"""Pins the profile-alignment oracle (oracle/pg_oracle.c), CPU only.

The reference's profile aligner is a stub (src/pgalign/myers/impl/
sequential.cpp:29-32 returns an empty alignment), so there is no reference
output to compare with: the oracle is pinned by what the mathematics fixes,
computed here from its output alone --
  - every row, gaps removed, is its sequence's first unpadded() codes, in
    ascending columns, and no column is empty;
  - every join's reported score is the sum-of-pairs score between the two
    children's members in the final alignment (gaps against gaps score 0), so
    every join kept its optimum through the later joins;
  - a join of two single sequences scores their Needleman-Wunsch optimum
    (textbook full matrix, and the reference's own sequential align() where
    it is built);
  - on tiny profiles, enumerating every way to merge the two children's
    alignments gives the same optimum;
and planted mistakes in the oracle break these pins.
"""
import itertools

import numpy as np
import pytest

from msa_amd import workloads as W
from oracle import museqa_ref as M
from tests import helpers as H

GAP = 25


def random_tree(rng, n):
    """A random binary guide tree in msa_nj_node form: join t is node n+t,
    children smaller id first, parents and levels filled."""
    nodes = np.zeros(max(2 * n - 1, 0), dtype=M.NJ_NODE)
    nodes["parent"] = M.NJ_UNDEFINED
    nodes["child"] = M.NJ_UNDEFINED
    active = list(range(n))
    for t in range(n - 1):
        a, b = rng.choice(len(active), 2, replace=False)
        x, y = sorted((active[a], active[b]))
        v = n + t
        nodes[v]["child"] = (x, y)
        nodes[x]["parent"] = v
        nodes[y]["parent"] = v
        nodes[v]["level"] = 1 + max(nodes[x]["level"], nodes[y]["level"])
        active = [u for u in active if u not in (x, y)] + [v]
    return nodes


def members(nodes, n, v):
    if v < n:
        return [v]
    c0, c1 = nodes[v]["child"]
    return members(nodes, n, int(c0)) + members(nodes, n, int(c1))


def sequence_codes(db, lengths):
    blocks, offsets, sizes = db
    return [W.unpack(blocks, int(o), int(s))[:int(L)] for o, s, L in zip(offsets, sizes, lengths)]


def gapped_rows(codes, cols, length):
    rows, at = [], 0
    for c in codes:
        row = np.full(length, GAP, dtype=np.int64)
        row[cols[at:at + len(c)]] = c
        rows.append(row)
        at += len(c)
    return rows


def sigma(T, p):
    s = np.full((26, 26), -p, dtype=np.int64)
    s[:25, :25] = T
    s[GAP, GAP] = 0
    return s


def sp_cross(rows_a, rows_b, sg):
    return int(sum(sg[a, b].sum() for a in rows_a for b in rows_b))


def induced(rows):
    """The alignment of some rows alone: columns where all of them hold a gap removed."""
    keep = np.any(np.array(rows) != GAP, axis=0)
    return [r[keep] for r in rows]


def brute_force_best(rows_a, rows_b, sg):
    """Every merge of two alignments (each column of A or B alone against
    gaps, or one of each together), scored directly."""
    la, lb = len(rows_a[0]), len(rows_b[0])
    gap_a = np.full(len(rows_a), GAP)
    gap_b = np.full(len(rows_b), GAP)
    A = np.array(rows_a).T
    B = np.array(rows_b).T
    best = None

    def walk(i, j, score):
        nonlocal best
        if i == la and j == lb:
            best = score if best is None else max(best, score)
            return
        if i < la and j < lb:
            walk(i + 1, j + 1, score + int(sg[np.ix_(A[i], B[j])].sum()))
        if i < la:
            walk(i + 1, j, score + int(sg[np.ix_(A[i], gap_b)].sum()))
        if j < lb:
            walk(i, j + 1, score + int(sg[np.ix_(gap_a, B[j])].sum()))

    walk(0, 0, 0)
    return best


def run(db, nodes, T, p, mistake=0):
    ora = M.PgOracle()
    lengths = ora.unpadded(*db)
    cols, length, scores = ora.align(*db, nodes, T, p, mistake=mistake)
    return lengths, cols, length, scores


def check_alignment(db, nodes, T, p, mistake=0, brute=False):
    """Every pin on one input; returns the first violated pin's name or None."""
    n = len(db[2])
    lengths, cols, length, scores = run(db, nodes, T, p, mistake)
    codes = sequence_codes(db, lengths)
    rows = gapped_rows(codes, cols, length)
    at = 0
    for c in codes:
        cc = cols[at:at + len(c)]
        if len(cc) and (np.any(np.diff(cc) <= 0) or cc[0] < 0 or cc[-1] >= length):
            return "columns"
        at += len(c)
    if n and length and np.any(np.all(np.array(rows) == GAP, axis=0)):
        return "empty column"
    sg = sigma(T, p)
    for v in range(n, 2 * n - 1):
        c0, c1 = (int(x) for x in nodes[v]["child"])
        ra = [rows[s] for s in members(nodes, n, c0)]
        rb = [rows[s] for s in members(nodes, n, c1)]
        if sp_cross(ra, rb, sg) != scores[v - n]:
            return "join score"
        if c0 < n and c1 < n:
            a = [int(x) for x in codes[c0]]
            b = [int(x) for x in codes[c1]]
            if M.textbook_nw(a, b, T, p) != scores[v - n]:
                return "pairwise"
        if brute:
            pa, pb = induced(ra), induced(rb)
            if brute_force_best(pa, pb, sg) != scores[v - n]:
                return "brute force"
    return None


def clean_db(rng, n, lo, hi, alphabet=24):
    return H.random_db(rng, n, lo, hi, alphabet=alphabet)


@pytest.mark.parametrize("seed", range(6))
def test_pins_on_random_trees(seed):
    rng = np.random.default_rng(seed)
    tables = H.tables()
    T, p = tables[["blosum62", "pam250", "default"][seed % 3]]
    n = int(rng.integers(2, 12))
    db = clean_db(rng, n, 0, 40)
    assert check_alignment(db, random_tree(rng, n), T, p) is None


@pytest.mark.parametrize("seed", range(8))
def test_brute_force_optimum_on_tiny_profiles(seed):
    rng = np.random.default_rng(100 + seed)
    T, p = H.tables()["blosum62"]
    n = int(rng.integers(2, 6))
    db = clean_db(rng, n, 0, 3)
    assert check_alignment(db, random_tree(rng, n), T, p, brute=True) is None


def test_neighbor_joining_tree_and_padding_codes():
    """The guide tree the phylogeny stage builds, and rows that hold internal
    padding codes ('*', which scores by the table's row 24)."""
    rng = np.random.default_rng(7)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 9, 1, 30, dirty=True)
    scores = H.oracle_scores(db, T, p)
    nodes = M.NjOracle().build(scores, 9)
    lengths, cols, length, jscores = run(db, nodes, T, p)
    assert len(cols) == lengths.sum()
    # Only the pairwise pin needs clean rows; the others hold as they are.
    n = 9
    codes = sequence_codes(db, lengths)
    rows = gapped_rows(codes, cols, length)
    sg = sigma(T, p)
    for v in range(n, 2 * n - 1):
        c0, c1 = (int(x) for x in nodes[v]["child"])
        assert sp_cross([rows[s] for s in members(nodes, n, c0)], [rows[s] for s in members(nodes, n, c1)],
                        sg) == jscores[v - n]


@pytest.mark.skipif(not M.ReferenceBuild.available(), reason="reference build (oracle/_ref) not present")
def test_two_sequences_score_the_reference_pairwise():
    """Two clean sequences: the one join scores what the reference's own
    sequential align() gives the pair."""
    rng = np.random.default_rng(8)
    ref = M.ReferenceBuild()
    for table in ("blosum62", "pam250", "default"):
        T, p = H.tables()[table]
        for _ in range(10):
            db = clean_db(rng, 2, 0, 60)
            nodes = random_tree(rng, 2)
            _, _, _, scores = run(db, nodes, T, p)
            exp = ref.align_list(db[0], db[1], db[2], np.array([[1, 0]], dtype=np.int32), T, p)
            assert scores[0] == int(exp[0])


def test_unpadded_length_quirk():
    """sequence::unpadded (src/sequence.hpp:110-119) counts padding only at
    positions 1 and 2 of the last block: 'A*C' keeps 'A*'."""
    ora = M.PgOracle()
    for text, want in [("", 0), ("A", 1), ("AC", 2), ("ACG", 3), ("ACGT", 4), ("A*C", 2), ("AC*", 2), ("*AC", 3),
                       ("A**", 1), ("ACGA*C", 5)]:
        db = H.strings_db([text])
        assert ora.unpadded(*db)[0] == want, text


def test_single_sequence_and_empty_rows():
    T, p = H.tables()["blosum62"]
    db = H.strings_db(["MKV"])
    nodes = np.zeros(1, dtype=M.NJ_NODE)
    nodes["parent"] = M.NJ_UNDEFINED
    nodes["child"] = M.NJ_UNDEFINED
    lengths, cols, length, scores = run(db, nodes, T, p)
    assert list(cols) == [0, 1, 2] and length == 3 and len(scores) == 0
    db = H.strings_db(["", "ACD", ""])
    nodes = random_tree(np.random.default_rng(1), 3)
    assert check_alignment(db, nodes, T, p) is None


@pytest.mark.parametrize("mistake,pin", [(1, "join score"), (2, "pairwise")])
def test_planted_mistakes_break_a_pin(mistake, pin):
    """Gap against gap scored -p (1) or a gap column charged per member, not
    per residue (2): some input breaks a pin."""
    T, p = H.tables()["blosum62"]
    found = None
    for seed in range(30):
        rng = np.random.default_rng(200 + seed)
        n = int(rng.integers(2, 6))
        db = clean_db(rng, n, 0, 4)
        found = check_alignment(db, random_tree(rng, n), T, p, mistake=mistake, brute=True)
        if found is not None:
            break
    assert found is not None, "planted mistake %d broke no pin" % mistake
    assert found in (pin, "brute force", "join score", "pairwise")


def test_tie_order_changes_the_alignment_not_the_score():
    """Ties taken left before diagonal (3) give other alignments of the same
    scores: nothing mathematical pins the tie rule, which is why the GPU is
    compared with the oracle column for column."""
    T, p = H.tables()["default"]
    differs = 0
    for seed in range(20):
        rng = np.random.default_rng(300 + seed)
        n = 2
        db = clean_db(rng, n, 1, 10, alphabet=4)
        nodes = random_tree(rng, n)
        _, c0, l0, s0 = run(db, nodes, T, p)
        _, c3, l3, s3 = run(db, nodes, T, p, mistake=3)
        assert list(s0) == list(s3)
        differs += int(l0 != l3 or not np.array_equal(c0, c3))
    assert differs > 0
