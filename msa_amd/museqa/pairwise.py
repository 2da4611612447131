"""museqa.pairwise (src/python/pairwise.pyx): ScoringTable, DistanceMatrix,
run and list.  Every named algorithm runs in the HIP kernels."""
from ._core import core

__all__ = ["ScoringTable", "DistanceMatrix", "run", "list"]

_list = list


class ScoringTable:
    def __init__(self, name=None, matrix=None, penalty=None):
        if matrix is not None:
            self._t = core.ScoringTable([[float(v) for v in row] for row in matrix], float(penalty))
        else:
            self._t = core.ScoringTable(name if name is not None else "default")

    def __getitem__(self, offset):
        x, y = (core.encode_unit(v) if isinstance(v, str) else int(v) for v in offset)
        return self._t.at(x, y)

    @staticmethod
    def list():
        return _list(core.ScoringTable.list())

    @property
    def penalty(self):
        return self._t.penalty

    @property
    def native(self):
        return self._t


class DistanceMatrix:
    def __init__(self, score=(), count=0, native=None):
        self._d = native if native is not None else core.DistanceMatrix([float(s) for s in score], int(count))

    def __getitem__(self, offset):
        return self._d.at(int(offset[0]), int(offset[1]))

    @property
    def count(self):
        return self._d.count

    def scores(self):
        return _list(self._d.raw())

    @property
    def native(self):
        return self._d


def list():  # noqa: A001 -- the reference's name
    return _list(core.algorithms())


def run(db, **kwargs):
    """Aligns every pair of `db`; algorithm may be a name or a Python callable
    taking (db, table=ScoringTable) as in the reference's binding."""
    table = kwargs.pop("table", ScoringTable())
    algorithm = kwargs.pop("algorithm", "default")
    table = table if isinstance(table, ScoringTable) else ScoringTable(table)
    if isinstance(algorithm, str):
        return DistanceMatrix(native=core.run(db.native, table.native, algorithm))
    return algorithm(db, table=table)
