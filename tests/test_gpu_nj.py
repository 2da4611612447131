"""GPU neighbor joining (msa_nj_build, msa_nj_build_device) against the oracle.

The oracle (oracle/nj_oracle.c) restates the reference's sequential neighbor
joining float for float and is pinned in test_nj_oracle.py; the GPU tree must
be the same node for node: parents, children, levels equal, branch lengths
bitwise equal (NaN where the oracle has NaN).
"""
import numpy as np
import pytest

from msa_amd import _native as N
from msa_amd import workloads as W
from oracle import museqa_ref as M
from tests import helpers as H

pytestmark = pytest.mark.gpu

ORACLE = M.NjOracle()


@pytest.fixture(scope="module", autouse=True)
def device():
    if N.device_count() < 1:
        pytest.skip("no GPU")
    N.check(N.lib().msa_device_select(0))


@pytest.mark.parametrize("variant", [N.NJ_SYMMETRIC, N.NJ_LINEAR])
@pytest.mark.parametrize("n", [2, 3, 4, 5, 17, 64, 250, 700, 1500])
def test_random_matrices(n, variant):
    rng = np.random.default_rng(n * 10 + variant)
    d = rng.normal(0, 100, n * (n - 1) // 2).astype(np.float32)
    H.assert_same_tree(N.nj_build(d, n, variant), ORACLE.build(d, n, variant))


@pytest.mark.parametrize("variant", [N.NJ_SYMMETRIC, N.NJ_LINEAR])
def test_ties(variant):
    """All distances equal: every Q ties at every iteration."""
    for n in (3, 7, 40):
        d = np.ones(n * (n - 1) // 2, dtype=np.float32)
        H.assert_same_tree(N.nj_build(d, n, variant), ORACLE.build(d, n, variant))


@pytest.mark.parametrize("variant", [N.NJ_SYMMETRIC, N.NJ_LINEAR])
def test_additive_trees(variant):
    for n, seed in [(6, 1), (33, 2), (120, 3)]:
        packed, splits, _ = M.additive_tree(np.random.default_rng(seed), n)
        got = N.nj_build(-packed, n, variant)
        H.assert_same_tree(got, ORACLE.build(-packed, n, variant))
        assert M.nj_splits(got, n) == splits


def test_small_inputs():
    assert len(N.nj_build(np.zeros(0, dtype=np.float32), 0)) == 0
    assert len(N.nj_build(np.zeros(0, dtype=np.float32), 1)) == 0
    got = N.nj_build(np.array([7.0], dtype=np.float32), 2)
    H.assert_same_tree(got, ORACLE.build(np.array([7.0], dtype=np.float32), 2))


@pytest.mark.parametrize("variant", [N.NJ_SYMMETRIC, N.NJ_LINEAR])
def test_pairwise_scores_chained_on_device(variant):
    """The phylogeny stage after the pairwise stage: int32 scores from a plan,
    left in device memory, feed msa_nj_build_device; the tree equals the
    oracle's on the same scores as floats, and msa_nj_build's on the host."""
    rng = np.random.default_rng(21)
    T, p = H.tables()["blosum62"]
    db = H.random_db(rng, 60, 20, 400)
    n = 60
    npairs = n * (n - 1) // 2
    plan = N.Plan(*db, T, p)
    d_scores = N.DeviceBuffer(4 * npairs)
    plan.execute(d_scores.ptr)
    plan.status()
    d_nodes = N.DeviceBuffer(N.NJ_NODE.itemsize * (2 * n - 1))
    N.nj_build_device(d_scores.ptr, n, variant, d_nodes.ptr)
    got = d_nodes.download(N.NJ_NODE, 2 * n - 1)
    scores = d_scores.download(np.int32, npairs).astype(np.float32)
    exp = ORACLE.build(scores, n, variant)
    H.assert_same_tree(got, exp)
    H.assert_same_tree(N.nj_build(scores, n, variant), exp)


def test_uniform_batch_512_tree():
    """The 512-sequence workload end to end: pairwise scores on the GPU, then
    the tree on the GPU, against the oracle's tree over the same scores."""
    w = W.uniform_batch(H.tables())
    scores = N.align(w.blocks, w.offsets, w.sizes, w.table, w.penalty).astype(np.float32)
    for variant in (N.NJ_SYMMETRIC, N.NJ_LINEAR):
        H.assert_same_tree(N.nj_build(scores, w.nseq, variant), ORACLE.build(scores, w.nseq, variant))
