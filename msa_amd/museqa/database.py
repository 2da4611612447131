"""museqa.database (src/python/database.pyx): Database."""
from collections import namedtuple

from ._core import core
from .sequence import Sequence

__all__ = ["Database"]


class Database:
    entry = namedtuple("entry", ["description", "contents"])

    def __init__(self, native=None):
        self._db = native if native is not None else core.Database()

    def __getitem__(self, key):
        if isinstance(key, bool) or not isinstance(key, (int, str, bytes)):
            raise TypeError("could not find sequence")
        if isinstance(key, bytes):
            key = key.decode()
        description, contents = self._db.at(key)
        return Database.entry(description, Sequence(contents))

    def add(self, target):
        if isinstance(target, dict):
            for item in target.items():
                self.add(item)
        elif isinstance(target, tuple):
            description, contents = target
            self._db.add(description, Sequence(contents).native)
        elif isinstance(target, list):
            for item in target:
                self.add(item)
        elif isinstance(target, Sequence):
            self._db.add_anonymous(target.native)
        elif isinstance(target, (str, bytes)):
            self._db.add_anonymous(Sequence(target).native)
        else:
            raise TypeError("could not add sequence to database")

    def merge(self, other):
        self._db.merge(other._db)

    def only(self, keys):
        if not all(type(k) is type(keys[0]) for k in keys):
            raise TypeError("cannot select sequences via different key types")
        if keys and isinstance(keys[0], int):
            return Database(self._db.only_offsets(keys))
        return Database(self._db.only_keys([k.decode() if isinstance(k, bytes) else k for k in keys]))

    @staticmethod
    def load(*filenames, device=False):
        """Loads FASTA files (src/io/loader/parser/fasta.cpp rules); with
        device=True the text is parsed on the GPU into the same database."""
        loader = core.Database.load_on_device if device else core.Database.load
        if len(filenames) == 1:
            return Database(loader(filenames[0]))
        result = core.Database()
        for name in filenames:
            result.merge(loader(name))
        return Database(result)

    @property
    def count(self):
        return self._db.count

    @property
    def native(self):
        return self._db
