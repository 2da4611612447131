"""museqa.phylogeny: neighbor joining over a pairwise DistanceMatrix
(src/phylogeny/phylogeny.cuh:114-134, src/phylogeny/phylogeny.cpp:26-41).
The reference ships no Python binding for this module; this one keeps the
pairwise module's shape (run, list).  Every named algorithm runs in the HIP
kernels."""
import numpy as np

from ._core import core

__all__ = ["GuideTree", "NODE", "UNDEFINED", "run", "list"]

_list = list

UNDEFINED = 0xFFFFFFFF

# One node: parent, children (smaller id first), branch length to the parent,
# level (src/phylogeny/tree.cuh, njoining/star.cuh:63-80).
NODE = np.dtype([("parent", "<u4"), ("child", "<u4", (2,)), ("distance", "<f4"), ("level", "<u4")])


class GuideTree:
    def __init__(self, native):
        self._t = native

    def __getitem__(self, ref):
        parent, child, distance, level = self._t.node(int(ref))
        return {"parent": parent, "child": child, "distance": distance, "level": level}

    def __len__(self):
        return self._t.count

    @property
    def leaves(self):
        return self._t.leaves

    @property
    def root(self):
        return self._t.root()

    def nodes(self):
        """Every node as a structured array of NODE (leaves first, root last)."""
        return np.frombuffer(self._t.raw(), dtype=NODE).copy()

    @property
    def native(self):
        return self._t


def list():  # noqa: A001 -- the pairwise module's name
    return _list(core.phylogeny_algorithms())


def run(matrix, count=None, algorithm="default"):
    """Joins the first `count` sequences of `matrix` (all of them by default)."""
    native = matrix.native
    count = native.count if count is None else int(count)
    return GuideTree(core.phylogeny_run(native, count, algorithm))
