"""museqa.sequence (src/python/sequence.pyx): Sequence and encode."""
from ._core import core

__all__ = ["Sequence", "encode"]


class Sequence:
    """An encoded sequence: 5-bit codes, three per block, padded with '*'."""

    padding = "*"

    def __init__(self, contents=b""):
        if isinstance(contents, core.Sequence):
            self._seq = contents
            return
        if isinstance(contents, str):
            contents = contents.encode("ascii")
        if not isinstance(contents, bytes):
            raise TypeError("unexpected argument type")
        self._seq = core.Sequence(contents)

    def __getitem__(self, offset):
        try:
            return self._seq[int(offset)]
        except (IndexError, ValueError) as e:
            raise IndexError(str(e)) from None

    def __str__(self):
        return str(self._seq)

    @property
    def length(self):
        return self._seq.length

    @property
    def native(self):
        return self._seq


def encode(sequence):
    """The padded decoded string the encoder produces for `sequence`."""
    units = [core.decode_unit(core.encode_unit(ch)) for ch in sequence]
    pad = len(sequence) % 3
    return "".join(units) + Sequence.padding * ((3 - pad) if pad else 0)
