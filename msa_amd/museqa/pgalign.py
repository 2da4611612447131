"""museqa.pgalign: the multiple alignment along a guide tree
(src/pgalign/pgalign.cuh:82-90, src/pgalign/pgalign.cpp:24-66).  The
reference ships no Python binding for this module; this one keeps the
pairwise module's shape (run, list).  Every named algorithm runs the GPU
progressive aligner (the reference's "myers" is a stub that returns an empty
alignment)."""
from ._core import core

__all__ = ["Alignment", "run", "list"]

_list = list


class Alignment:
    """One gapped row per sequence, in database order: the column of each
    residue (src/pgalign/sequence.cuh) and the row as text with '-' gaps."""

    def __init__(self, native):
        self._a = native

    def __len__(self):
        return self._a.count

    @property
    def length(self):
        return self._a.length

    def columns(self, row):
        return self._a.row(int(row))[0]

    def __getitem__(self, row):
        return self._a.row(int(row))[1]

    def at(self, row, column):
        return self._a.at(int(row), int(column))

    def slice(self, displ, size):
        return Alignment(self._a.slice(int(displ), int(size)))

    @property
    def native(self):
        return self._a


def list():  # noqa: A001 -- the pairwise module's name
    return _list(core.pgalign_algorithms())


def run(db, tree, count=None, table=None, algorithm="default"):
    """Aligns the first `count` sequences of `db` (all the tree's leaves by
    default) along `tree`, scoring with `table` (a pairwise ScoringTable or a
    table name; the "default" table when None)."""
    from . import pairwise
    if table is None:
        table = "default"
    if isinstance(table, str):
        table = pairwise.ScoringTable(table)
    count = tree.leaves if count is None else int(count)
    return Alignment(core.pgalign_run(db.native, tree.native, count, table.native, algorithm))
