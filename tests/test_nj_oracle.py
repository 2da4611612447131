"""The neighbor-joining oracle (oracle/nj_oracle.c), pinned on the CPU.

The reference's sequential neighbor joining
(src/phylogeny/njoining/impl/sequential.cpp) cannot be built here (its matrix
moves live in a .cu file and its node bookkeeping in the MPI layer), and the
reference's tests hold no phylogeny case.  The restatement is pinned by what
the mathematics fixes:

  - additive trees: for a path-length metric D of a binary tree, neighbor
    joining recovers the tree and its edge lengths (Saitou & Nei 1987; Studier
    & Keppler 1988).  The reference joins the pair of largest Q over
    similarities, so it is fed S = -D: Q_S = -Q_D and every branch length
    comes out negated;
  - ties: with all distances equal every Q of every iteration ties, so the
    reference's strict `>` over pairs (i > j) in row-major order always joins
    logical pair (1, 0); the tree then follows from the index bookkeeping
    alone, which differs between the symmetric and the linear variant;
  - structure: 2n-1 nodes, parents numbered in join order, children by id,
    levels.
"""
import numpy as np
import pytest

from oracle import museqa_ref as M

ORACLE = M.NjOracle()


def branch(nodes, v):
    return float(nodes[v]["distance"])


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n,seed", [(4, 1), (5, 2), (6, 3), (9, 4), (16, 5), (33, 6), (60, 7)])
def test_additive_tree_is_recovered(n, seed, variant):
    rng = np.random.default_rng(seed)
    packed, splits, side_len = M.additive_tree(rng, n)
    nodes = ORACLE.build(-packed, n, variant)
    assert M.nj_splits(nodes, n) == splits
    # Every edge below the root's two children has its length, negated.
    root = 2 * n - 2
    below_root = set(int(c) for c in nodes[root]["child"])
    children = {v: [int(c) for c in nodes[v]["child"]] for v in range(n, 2 * n - 1)}
    for v in range(2 * n - 2):
        if v in below_root:
            continue
        leaves = leaf_set(children, v, n)
        key = frozenset(range(n)) - leaves if 0 in leaves else leaves
        assert abs(-branch(nodes, v) - side_len[key]) < 1e-3 * max(1.0, side_len[key]), (v, key)


def leaf_set(children, v, n):
    if v < n:
        return frozenset([v])
    return frozenset().union(*[leaf_set(children, c, n) for c in children[v]])


def caterpillar(n, variant):
    """The tree when every join takes logical pair (1, 0): simulate only the
    reference's index bookkeeping (sequential.cpp:188-262)."""
    labels = list(range(n))
    parents = []
    for t in range(n - 1):
        x, y = 1, 0
        a, b = labels[x], labels[y]
        parents.append(tuple(sorted((a, b))))
        labels[x] = n + t
        if variant == 0:                    # swap y with 0, remove 0
            labels[0], labels[y] = labels[y], labels[0]
            labels = labels[1:]
        else:                               # swap y with the last, remove it
            labels[-1], labels[y] = labels[y], labels[-1]
            labels = labels[:-1]
    return parents


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n", [2, 3, 4, 7, 12])
def test_ties_join_the_first_pair(n, variant):
    nodes = ORACLE.build(np.ones(n * (n - 1) // 2, dtype=np.float32), n, variant)
    got = [tuple(int(c) for c in nodes[n + t]["child"]) for t in range(n - 1)]
    assert got == caterpillar(n, variant)


def test_variants_differ_on_ties():
    assert caterpillar(6, 0) != caterpillar(6, 1)


@pytest.mark.parametrize("variant", [0, 1])
def test_tree_structure(variant):
    rng = np.random.default_rng(3)
    n = 40
    nodes = ORACLE.build(rng.normal(0, 100, n * (n - 1) // 2).astype(np.float32), n, variant)
    assert len(nodes) == 2 * n - 1
    root = 2 * n - 2
    assert nodes[root]["parent"] == M.NJ_UNDEFINED
    for v in range(n):
        assert nodes[v]["level"] == 0 and tuple(nodes[v]["child"]) == (M.NJ_UNDEFINED, M.NJ_UNDEFINED)
    for v in range(2 * n - 2):
        assert v < nodes[v]["parent"] <= root
    for v in range(n, 2 * n - 1):
        c1, c2 = (int(c) for c in nodes[v]["child"])
        assert c1 < c2 < v
        assert nodes[c1]["parent"] == v and nodes[c2]["parent"] == v
        assert nodes[v]["level"] == max(nodes[c1]["level"], nodes[c2]["level"]) + 1


def test_last_join_divides_by_zero_like_the_reference():
    """At two OTUs the reference's branch length divides by 2 * (count - 2) = 0
    (sequential.cpp:141): the root's children get +-inf (NaN when the row sums
    are equal)."""
    nodes = ORACLE.build(np.array([3.0, 1.0, 2.0], dtype=np.float32), 3, 0)
    root = 4
    d = [branch(nodes, int(c)) for c in nodes[root]["child"]]
    assert all(np.isinf(x) or np.isnan(x) for x in d)


def test_small_inputs():
    assert len(ORACLE.build(np.zeros(0, dtype=np.float32), 0)) == 0
    assert len(ORACLE.build(np.zeros(0, dtype=np.float32), 1)) == 0
    nodes = ORACLE.build(np.array([5.0], dtype=np.float32), 2)
    assert tuple(nodes[2]["child"]) == (0, 1)


def pins_hold(mistake, variant):
    """True when every pin above holds for the oracle with `mistake` planted."""
    for n, seed in [(6, 3), (16, 5), (33, 6)]:
        rng = np.random.default_rng(seed)
        packed, splits, side_len = M.additive_tree(rng, n)
        nodes = ORACLE.build(-packed, n, variant, mistake=mistake)
        if M.nj_splits(nodes, n) != splits:
            return False
        root = 2 * n - 2
        children = {v: [int(c) for c in nodes[v]["child"]] for v in range(n, 2 * n - 1)}
        for v in range(2 * n - 2):
            if v in set(int(c) for c in nodes[root]["child"]):
                continue
            leaves = leaf_set(children, v, n)
            key = frozenset(range(n)) - leaves if 0 in leaves else leaves
            if not abs(-branch(nodes, v) - side_len[key]) < 1e-3 * max(1.0, side_len[key]):
                return False
    for n in (7, 12):
        nodes = ORACLE.build(np.ones(n * (n - 1) // 2, dtype=np.float32), n, variant, mistake=mistake)
        if [tuple(int(c) for c in nodes[n + t]["child"]) for t in range(n - 1)] != caterpillar(n, variant):
            return False
    return True


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("mistake", [1, 2, 3, 4])
def test_planted_mistakes_fail_a_pin(mistake, variant):
    """Q with (count-1), the smallest Q, an update without .5, the other
    variant's removal: each breaks a pin; the unplanted oracle keeps them all."""
    assert pins_hold(0, variant)
    assert not pins_hold(mistake, variant)
