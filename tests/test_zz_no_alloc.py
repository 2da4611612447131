"""msa_nw_plain_noalloc: the plain entry point with its workspace in a fixed
device array (no hipMalloc), against msa_nw_plain through the C ABI.

Two calls in a row, the second with a smaller input (the array keeps the
first call's contents), each compared pair for pair with msa_nw_plain on the
same device buffers and with the oracle; then the documented limits."""
import numpy as np
import pytest

from msa_amd import _native as N
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if N.device_count() < 1:
        pytest.skip("no GPU")
    N.check(N.lib().msa_device_select(0))


class DeviceDb:
    def __init__(self, db):
        blocks, offsets, sizes = db
        self.n = len(sizes)
        self.max_blocks = int(sizes.max()) if len(sizes) else 0
        self.bufs = []
        self.blocks = self._up(blocks)
        self.offsets = self._up(offsets)
        self.sizes = self._up(sizes)

    def _up(self, a):
        b = N.DeviceBuffer(max(a.nbytes, 4))
        if a.nbytes:
            b.upload(a)
        self.bufs.append(b)
        return b.ptr

    def close(self):
        for b in self.bufs:
            b.close()


def run(entry, ddb, T, p, pairs, npairs, d_out):
    table = np.ascontiguousarray(T, dtype=np.int32).reshape(625)
    N.check(entry(ddb.blocks, ddb.offsets, ddb.sizes, ddb.n, ddb.max_blocks, pairs, npairs, table.ctypes.data, p,
                  d_out.ptr, None))
    N.check(N.lib().msa_device_sync())
    return d_out.download(np.int32, npairs)


def test_noalloc_equals_plain_entry_on_two_calls():
    lib = N.lib()
    rng = np.random.default_rng(91)
    T, p = H.tables()["blosum62"]
    # First call: 48 sequences up to 1,400 residues (several strips), padding
    # codes inside, all pairs in reference order.
    db1 = H.random_db(rng, 48, 0, 1400, dirty=True)
    d1 = DeviceDb(db1)
    n1 = 48 * 47 // 2
    out_a, out_b = N.DeviceBuffer(4 * n1), N.DeviceBuffer(4 * n1)
    # Second call: fewer, shorter sequences and an explicit pair list, under
    # another table.
    db2 = H.random_db(rng, 12, 0, 300, dirty=True)
    d2 = DeviceDb(db2)
    pairs = rng.integers(0, 12, (90, 2)).astype(np.int32)
    d_pairs = N.DeviceBuffer(pairs.nbytes)
    d_pairs.upload(pairs)
    T2, p2 = H.tables()["pam250"]
    try:
        got1 = run(lib.msa_nw_plain_noalloc, d1, T, p, None, n1, out_a)
        exp1 = run(lib.msa_nw_plain, d1, T, p, None, n1, out_b)
        assert np.array_equal(got1, exp1)
        assert np.array_equal(got1.astype(np.float32), H.oracle_scores(db1, T, p))
        N.check(lib.msa_memset(out_a.ptr, 0xAB, 4 * n1))
        got2 = run(lib.msa_nw_plain_noalloc, d2, T2, p2, d_pairs.ptr, len(pairs), out_a)
        exp2 = run(lib.msa_nw_plain, d2, T2, p2, d_pairs.ptr, len(pairs), out_b)
        assert np.array_equal(got2, exp2)
        assert np.array_equal(got2.astype(np.float32), H.oracle_scores(db2, T2, p2, pairs=pairs))
    finally:
        for b in (out_a, out_b, d_pairs):
            b.close()
        d1.close()
        d2.close()


def test_noalloc_refuses_sizes_beyond_its_array():
    """More sequences than the SeqInfo region holds, or a sequence whose
    strip-boundary line alone exceeds the line region: MSA_ERR_TOO_LARGE,
    before any device work."""
    lib = N.lib()
    T, p = H.tables()["blosum62"]
    table = np.ascontiguousarray(T, dtype=np.int32).reshape(625)
    d = DeviceDb(H.strings_db(["ACGT", "ACG", "AC"]))
    out = N.DeviceBuffer(64)
    try:
        rc = lib.msa_nw_plain_noalloc(d.blocks, d.offsets, d.sizes, 200000, 2, d.blocks, 1, table.ctypes.data, p,
                                      out.ptr, None)
        assert rc == N.MSA_ERR_TOO_LARGE
        rc = lib.msa_nw_plain_noalloc(d.blocks, d.offsets, d.sizes, 3, 6 * 1000 * 1000, None, 3,
                                      table.ctypes.data, p, out.ptr, None)
        assert rc == N.MSA_ERR_TOO_LARGE
        rc = lib.msa_nw_plain_noalloc(d.blocks, d.offsets, d.sizes, 3, 2, None, 4, table.ctypes.data, p, out.ptr,
                                      None)
        assert rc == N.MSA_ERR_INVALID_ARGUMENT
    finally:
        out.close()
        d.close()
