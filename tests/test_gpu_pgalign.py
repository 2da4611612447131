"""The profile-alignment stage on the GPU (msa_pg_align) against the oracle
(oracle/pg_oracle.c, pinned in test_pg_oracle.py): the same columns for every
residue, the same length and the same optimal score for every join, on random
trees and on the neighbor-joining tree of the pairwise scores; at sizes the
oracle does not reach, the properties it pins."""
import numpy as np
import pytest

from msa_amd import _native as N
from msa_amd import workloads as W
from oracle import museqa_ref as M
from tests import helpers as H
from tests.test_pg_oracle import gapped_rows, random_tree, sequence_codes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if N.device_count() < 1:
        pytest.skip("no GPU")
    N.check(N.lib().msa_device_select(0))


def assert_same(db, nodes, T, p):
    got = N.pg_align(*db, nodes, T, p)
    exp = M.PgOracle().align(*db, nodes, T, p)
    assert got[1] == exp[1], "length gpu %d oracle %d" % (got[1], exp[1])
    assert np.array_equal(got[2], exp[2]), "join scores differ"
    bad = np.nonzero(got[0] != exp[0])[0]
    assert len(bad) == 0, "%d of %d columns differ, first at %s" % (len(bad), len(exp[0]), bad[:5])
    return got


@pytest.mark.parametrize("seed", range(8))
def test_random_trees_match_the_oracle(seed):
    rng = np.random.default_rng(seed)
    table = ["blosum62", "pam250", "default", "blosum80"][seed % 4]
    T, p = H.tables()[table]
    n = int(rng.integers(2, 40))
    db = H.random_db(rng, n, 0, 120, dirty=seed % 2 == 1)
    assert_same(db, random_tree(rng, n), T, p)


def test_neighbor_joining_tree_of_gpu_scores():
    """The pipeline as it runs: pairwise scores on the GPU, the NJ tree on the
    GPU, the alignment along it."""
    w = W.uniform_small(H.tables())
    db = (w.blocks, w.offsets, w.sizes)
    scores = N.align(*db, w.table, w.penalty)
    nodes = N.nj_build(scores.astype(np.float32), w.nseq)
    cols, length, jscores = assert_same(db, nodes, w.table, w.penalty)
    assert length >= int(w.lengths.max())


def test_dna_caterpillar_and_long_rows():
    """A chain-shaped tree (every join adds one sequence) over DNA rows of
    1,000-3,000 residues: deep trees and wide joins."""
    rng = np.random.default_rng(11)
    T, p = H.tables()["pam250"]
    n = 12
    db = H.random_db(rng, n, 1000, 3000, alphabet=4)
    nodes = np.zeros(2 * n - 1, dtype=M.NJ_NODE)
    nodes["parent"] = M.NJ_UNDEFINED
    nodes["child"] = M.NJ_UNDEFINED
    prev = 0
    for t in range(n - 1):
        v = n + t
        nodes[v]["child"] = (prev, t + 1) if prev < t + 1 else (t + 1, prev)
        nodes[v]["level"] = t + 1
        prev = v
    assert_same(db, nodes, T, p)


def test_edge_cases():
    T, p = H.tables()["blosum62"]
    one = np.zeros(1, dtype=M.NJ_NODE)
    one["parent"] = M.NJ_UNDEFINED
    one["child"] = M.NJ_UNDEFINED
    cols, length, scores = N.pg_align(*H.strings_db(["MKVL"]), one, T, p)
    assert list(cols) == [0, 1, 2, 3] and length == 4 and len(scores) == 0
    rng = np.random.default_rng(3)
    assert_same(H.strings_db(["", "ACD", "", "A*C", "W"]), random_tree(rng, 5), T, p)
    assert_same(H.strings_db(["", ""]), random_tree(rng, 2), T, p)
    with pytest.raises(N.MsaError) as e:
        N.pg_align(*H.strings_db(["AC-GT", "ACGT"]), random_tree(rng, 2), T, p)
    assert e.value.status == N.MSA_ERR_GAP_CODE
    bad = random_tree(rng, 3)
    bad[4]["child"] = (4, 1)
    with pytest.raises(N.MsaError) as e:
        N.pg_align(*H.strings_db(["A", "C", "D"]), bad, T, p)
    assert e.value.status == N.MSA_ERR_INVALID_ARGUMENT


def test_serving_sized_job_properties():
    """A job of 200 protein sequences of 100-600 residues along its NJ tree:
    every row holds its sequence in ascending columns, no column is empty, and
    every join of two leaves induces exactly the pairwise stage's score."""
    tables = H.tables()
    T, p = tables["blosum62"]
    rng = np.random.default_rng(5)
    n = 200
    lengths = rng.integers(100, 600, n)
    codes = [W.AMINO20_CODES[rng.integers(0, 20, int(L))] for L in lengths]
    db = W.pack(codes)
    scores = N.align(*db, T, p)
    nodes = N.nj_build(scores.astype(np.float32), n)
    cols, length, jscores = N.pg_align(*db, nodes, T, p)
    rows = gapped_rows(codes, cols, length)
    at = 0
    for c in codes:
        cc = cols[at:at + len(c)]
        assert np.all(np.diff(cc) > 0) and cc[0] >= 0 and cc[-1] < length
        at += len(c)
    assert not np.any(np.all(np.array(rows) == 25, axis=0))
    sg = np.full((26, 26), -p, dtype=np.int64)
    sg[:25, :25] = T
    sg[25, 25] = 0
    leaf_joins = 0
    for v in range(n, 2 * n - 1):
        a, b = (int(x) for x in nodes[v]["child"])
        if a < n and b < n:
            i, j = max(a, b), min(a, b)
            assert int(sg[rows[a], rows[b]].sum()) == int(scores[i * (i - 1) // 2 + j]) == int(jscores[v - n])
            leaf_joins += 1
    assert leaf_joins > 0
