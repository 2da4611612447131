"""ORACLE -- test infrastructure, not part of the product.

Python restatements of the reference's host-side behaviour on the pairwise
path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module, and only as the checker.

Contents and what each follows:
  encode_char / encode / decode ......... src/encoder.cpp:27-118 (letter map,
                                          3 codes per uint16 block at shifts
                                          1/6/11, bit 0 = block holds padding)
  parse_fasta ........................... src/io/loader/parser/fasta.cpp:26-70
  needleman ............................. src/python/algorithm/pairwise/needleman.py:19-63
                                          (pure-Python loops: small inputs only)
  pair_order ............................ src/pairwise/pairwise.cpp:73-85
  textbook_nw ........................... independent textbook global alignment
                                          with a linear gap (full matrix, int64)
  CpuOracle ............................. ctypes wrapper of nw_oracle.c, the C
                                          restatement of sequential.cpp:29-95
  NjOracle .............................. ctypes wrapper of nj_oracle.c, the C
                                          restatement of the sequential
                                          neighbor joining
                                          (src/phylogeny/njoining/impl/
                                          sequential.cpp)
  tree_splits / additive_tree ........... textbook helpers that pin NjOracle:
                                          the bipartitions of a tree and the
                                          path-length metric of a random tree
  FastaOracle ........................... ctypes wrapper of fasta_oracle.c, the
                                          C restatement of fasta.cpp:26-70 +
                                          encoder.cpp:76-98 + the database
                                          packing, over bytes in memory
  ReferenceFasta ........................ ctypes wrapper of oracle/_ref/
                                          libref_fasta.so: the reference's own
                                          fasta.cpp, encoder.cpp, database.cpp
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

END = 0x18
GAP = 0x19
SHIFT = (1, 6, 11)

# src/encoder.cpp:31-34, indexed by upper - 'A'
_ENCODE = (0x00, 0x14, 0x01, 0x06, 0x08, 0x0E, 0x03, 0x09, 0x0A, 0x15, 0x0C, 0x0B, 0x0D,
           0x05, 0x17, 0x0F, 0x07, 0x04, 0x10, 0x02, 0x17, 0x13, 0x11, 0x17, 0x12, 0x16)
# src/encoder.cpp:40-43
_DECODE = "ACTGRNDQEHILKMFPSWYVBJZX*-"


def encode_char(ch):
    """encoder::encode(char), src/encoder.cpp:50-61."""
    upper = ch.upper() if "a" <= ch <= "z" else ch
    if "A" <= upper <= "Z":
        return _ENCODE[ord(upper) - ord("A")]
    if ch == "-":
        return GAP
    return END


def encode(text):
    """encoder::encode(ptr, size), src/encoder.cpp:76-98 -> list of uint16 blocks."""
    size = len(text)
    blocks = []
    n = 0
    while n < size:
        block = 1 if n + 3 > size else 0
        for k in range(3):
            code = encode_char(text[n]) if n < size else END
            block |= code << SHIFT[k]
            n += 1
        blocks.append(block & 0xFFFF)
    return blocks


def access(blocks, offset):
    """encoder::access, src/encoder.hpp:54-70."""
    return (blocks[offset // 3] >> SHIFT[offset % 3]) & 0x1F


def decode(blocks):
    """encoder::decode(buffer), src/encoder.cpp:105-118 (padded string)."""
    return "".join(_DECODE[access(blocks, i)] for i in range(3 * len(blocks)))


def parse_fasta(path):
    """io::parser::fasta, src/io/loader/parser/fasta.cpp:26-70.

    A record starts at a '>' line and takes the following lines up to the
    next '>' or the first empty line.  Returns [(description, sequence)].
    """
    with open(path, "r", newline="") as f:
        text = f.read()
    lines = text.split("\n")
    has_final_newline = text.endswith("\n")
    if has_final_newline:
        lines = lines[:-1]
    out = []
    k = 0
    while k < len(lines):
        while k < len(lines) and (len(lines[k]) < 1 or lines[k][0] != ">"):
            k += 1
        if k >= len(lines):
            break
        description = lines[k][1:]
        k += 1
        contents = []
        while k < len(lines) and not lines[k].startswith(">") and len(lines[k]) > 0:
            contents.append(lines[k])
            k += 1
        if k < len(lines) and len(lines[k]) == 0:
            k += 1
        out.append((description, "".join(contents)))
    return out


def pair_order(n):
    """algorithm::generate, src/pairwise/pairwise.cpp:73-85."""
    return [(i, j) for i in range(1, n) for j in range(i)]


def load_tables():
    with open(os.path.join(HERE, "tables.json")) as f:
        data = json.load(f)
    return {t["name"]: (np.array(t["matrix"], dtype=np.int64), t["penalty"]) for t in data["tables"]}


def needleman(seqs, table, penalty):
    """needleman.py:19-63 over block lists; returns float scores in pair order.

    `seqs` are block lists, `table` a 25x25 nested list or array, `penalty`
    the gap penalty.  Padding is code 24, as Sequence.padding ('*') is.
    """
    def align(one, two):
        if 3 * len(one) <= 3 * len(two):
            one, two = two, one
        n1, n2 = 3 * len(one), 3 * len(two)
        line = [float(i * -penalty) for i in range(n2 + 1)]
        for i in range(n1):
            a = access(one, i)
            if a == END:
                break
            done = line[0]
            line[0] = float((i + 1) * -penalty)
            for j in range(1, n2 + 1):
                value = line[j - 1]
                b = access(two, j - 1)
                if b != END:
                    insertd = value - penalty
                    removed = line[j] - penalty
                    matched = done + float(table[a][b])
                    value = max(insertd, removed, matched)
                done = line[j]
                line[j] = value
        return line[-1]

    return [align(seqs[i], seqs[j]) for i, j in pair_order(len(seqs))]


def textbook_nw(a, b, table, penalty):
    """Textbook global alignment score, linear gap, full DP in int64.

    H[i][0] = -i*p, H[0][j] = -j*p,
    H[i][j] = max(H[i-1][j-1] + T[a_i][b_j], H[i-1][j] - p, H[i][j-1] - p).
    Each row is computed with a prefix maximum:  H[i][j] = max_k<=j (t_k - (j-k)p).
    """
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    table = np.asarray(table, dtype=np.int64)
    p = int(penalty)
    m = len(b)
    jp = np.arange(m + 1, dtype=np.int64) * p
    prev = -jp.copy()
    for i in range(1, len(a) + 1):
        t = np.empty(m + 1, dtype=np.int64)
        t[0] = -i * p
        if m:
            t[1:] = np.maximum(prev[:-1] + table[a[i - 1], b], prev[1:] - p)
        prev = np.maximum.accumulate(t + jp) - jp
    return int(prev[m])


class CpuOracle:
    """ctypes wrapper of oracle/_build/liboracle.so (nw_oracle.c)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        self.lib = ctypes.CDLL(path)
        self.lib.oracle_align_range.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_float,
            ctypes.c_void_p, ctypes.c_int]
        self.lib.oracle_align_range.restype = ctypes.c_int

    def align(self, blocks, offsets, nblocks, table625, penalty, pairs=None,
              first=0, count=None, workers=1):
        """Scores of pairs [first, first+count) as float32 (reference order)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nblocks = np.ascontiguousarray(nblocks, dtype=np.int64)
        table = np.ascontiguousarray(np.asarray(table625, dtype=np.float32).reshape(625))
        n = len(nblocks)
        if pairs is not None:
            pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
            total = len(pairs)
        else:
            total = n * (n - 1) // 2
        if count is None:
            count = total - first
        out = np.zeros(count, dtype=np.float32)
        if count == 0:
            return out
        rc = self.lib.oracle_align_range(
            blocks.ctypes.data if len(blocks) else None, offsets.ctypes.data, nblocks.ctypes.data,
            pairs.ctypes.data if pairs is not None else None, first, count,
            table.ctypes.data, float(penalty), out.ctypes.data, int(workers))
        if rc != 0:
            raise RuntimeError("oracle failed")
        return out


class ReferenceBuild:
    """ctypes wrapper of oracle/_ref/libref_sequential.so: the reference's own
    sequential.cpp + encoder.cpp compiled from its sources (oracle/Makefile,
    target `ref`).  Present only where it was built."""

    PATH = os.path.join(HERE, "_ref", "libref_sequential.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self):
        self.lib = ctypes.CDLL(self.PATH)
        self.lib.ref_align_list.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
            ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_int]
        self.lib.ref_align_list.restype = ctypes.c_int
        self.lib.ref_align_pair.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.c_void_p, ctypes.c_float]
        self.lib.ref_align_pair.restype = ctypes.c_float
        self.lib.ref_encode.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        self.lib.ref_encode.restype = ctypes.c_size_t

    def encode(self, text):
        raw = text.encode("latin-1")
        out = np.zeros(len(raw) // 3 + 2, dtype=np.uint16)
        n = self.lib.ref_encode(raw, len(raw), out.ctypes.data, len(out))
        return [int(v) for v in out[:n]]

    def align_list(self, blocks, offsets, nblocks, pairs, table625, penalty, workers=1):
        """Scores of an explicit pair list through the reference's own align()
        (sequential.cpp:77-95) over a museqa::database of every sequence, which
        chooses each pair's orientation itself."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nblocks = np.ascontiguousarray(nblocks, dtype=np.int64)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        table = np.ascontiguousarray(np.asarray(table625, dtype=np.float32).reshape(625))
        out = np.zeros(len(pairs), dtype=np.float32)
        if len(pairs):
            rc = self.lib.ref_align_list(blocks.ctypes.data if len(blocks) else None, offsets.ctypes.data,
                                         nblocks.ctypes.data, len(nblocks), pairs.ctypes.data, len(pairs),
                                         table.ctypes.data, float(penalty), out.ctypes.data, int(workers))
            if rc != 0:
                raise ValueError("pair list names a sequence outside the database")
        return out


# msa_nj_node / nj_node (oracle/nj_oracle.c)
NJ_NODE = np.dtype([("parent", np.uint32), ("child", np.uint32, (2,)), ("distance", np.float32),
                    ("level", np.uint32)])
NJ_UNDEFINED = 0xFFFFFFFF


class NjOracle:
    """ctypes wrapper of nj_oracle.c (in oracle/_build/liboracle.so)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        self.lib = ctypes.CDLL(path)
        self.lib.nj_oracle_planted.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_void_p]
        self.lib.nj_oracle_planted.restype = ctypes.c_int

    def build(self, distances, n, variant=0, mistake=0):
        """Nodes (NJ_NODE records, 2n-1) of the tree over n(n-1)/2 float
        distances in pair order; variant 0 = symmetric, 1 = linear.  A non-zero
        `mistake` plants an error (tests of the pins only)."""
        d = np.ascontiguousarray(distances, dtype=np.float32)
        assert d.size == n * (n - 1) // 2
        nodes = np.zeros(2 * n - 1 if n >= 2 else 0, dtype=NJ_NODE)
        if n >= 2 and self.lib.nj_oracle_planted(d.ctypes.data, n, variant, mistake, nodes.ctypes.data) != 0:
            raise RuntimeError("oracle failed")
        return nodes


class PgOracle:
    """ctypes wrapper of pg_oracle.c (in oracle/_build/liboracle.so): the
    progressive profile alignment along a guide tree."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        self.lib = ctypes.CDLL(path)
        self.lib.pg_oracle_planted.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int,
                                               ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]
        self.lib.pg_oracle_planted.restype = ctypes.c_int
        self.lib.pg_unpadded.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        self.lib.pg_unpadded.restype = ctypes.c_int64

    def unpadded(self, blocks, offsets, sizes):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        return np.array([self.lib.pg_unpadded(blocks[int(o):].ctypes.data if int(s) else None, int(s))
                         for o, s in zip(offsets, sizes)], dtype=np.int64)

    def align(self, blocks, offsets, sizes, nodes, table625, penalty, mistake=0):
        """(columns per residue in database order, length, join scores).  A
        non-zero `mistake` plants an error (tests of the pins only)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        nodes = np.ascontiguousarray(nodes, dtype=NJ_NODE)
        table = np.ascontiguousarray(np.asarray(table625).reshape(625), dtype=np.int32)
        n = len(sizes)
        total = int(self.unpadded(blocks, offsets, sizes).sum()) if n else 0
        cols = np.zeros(max(total, 1), dtype=np.int32)
        scores = np.zeros(max(n - 1, 1), dtype=np.int64)
        length = ctypes.c_int64(0)
        rc = self.lib.pg_oracle_planted(blocks.ctypes.data if len(blocks) else None, offsets.ctypes.data,
                                        sizes.ctypes.data, n, nodes.ctypes.data if len(nodes) else None,
                                        table.ctypes.data, int(penalty), int(mistake), cols.ctypes.data,
                                        ctypes.byref(length), scores.ctypes.data)
        if rc != 0:
            raise ValueError("invalid guide tree or gap code")
        return cols[:total], length.value, scores[:max(n - 1, 0)]


def tree_splits(children, n):
    """Bipartitions of the leaves 0..n-1 of a tree, each as the frozenset of
    the side that does not hold leaf 0; `children` maps an internal node to
    its children.  Trivial splits (one leaf, or all but one) are left out."""
    below = {}

    def leaves(v):
        if v < n:
            return frozenset([v])
        if v not in below:
            below[v] = frozenset().union(*[leaves(c) for c in children[v]])
        return below[v]

    full = frozenset(range(n))
    out = set()
    for v in children:
        side = leaves(v)
        if 0 in side:
            side = full - side
        if 1 < len(side) < n - 1:
            out.add(side)
    return out


def nj_splits(nodes, n):
    """tree_splits of an NJ_NODE tree."""
    children = {v: [int(c) for c in nodes[v]["child"]] for v in range(n, len(nodes))}
    return tree_splits(children, n)


def additive_tree(rng, n, lo=1, hi=20):
    """A random unrooted binary tree on leaves 0..n-1 with integer edge
    lengths in [lo, hi]; returns (packed path-length distances in pair order
    as float32, its splits, {frozenset side: edge length} for every edge)."""
    # Random topology by sequential insertion: each new leaf splits an edge.
    edges = {(0, n), (1, n), (2, n)}        # node n is the first internal node
    nxt = n + 1
    for leaf in range(3, n):
        a, b = sorted(edges)[int(rng.integers(0, len(edges)))]
        edges.remove((a, b))
        mid = nxt
        nxt += 1
        edges |= {(a, mid), (b, mid), (leaf, mid)}
    length = {e: int(rng.integers(lo, hi + 1)) for e in edges}
    adj = {}
    for (a, b), w in length.items():
        adj.setdefault(a, []).append((b, w))
        adj.setdefault(b, []).append((a, w))
    dist = np.zeros((n, n), dtype=np.int64)
    side_len = {}
    for src in range(n):
        stack, seen = [(src, 0)], {src}
        while stack:
            v, d = stack.pop()
            if v < n:
                dist[src, v] = d
            for u, w in adj[v]:
                if u not in seen:
                    seen.add(u)
                    stack.append((u, d + w))
    # the split of each edge: leaves on the far side of (a, b) from a
    for (a, b), w in length.items():
        stack, seen, side = [b], {a, b}, set()
        while stack:
            v = stack.pop()
            if v < n:
                side.add(v)
            for u, _ in adj[v]:
                if u not in seen:
                    seen.add(u)
                    stack.append(u)
        side = frozenset(side)
        if 0 in side:
            side = frozenset(range(n)) - side
        side_len[side] = w
    packed = np.array([dist[i, j] for i in range(1, n) for j in range(i)], dtype=np.float32)
    splits = {s for s in side_len if 1 < len(s) < n - 1}
    return packed, splits, side_len


class FastaOracle:
    """ctypes wrapper of fasta_oracle.c (in oracle/_build/liboracle.so)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_build", "liboracle.so")
        self.lib = ctypes.CDLL(path)
        self.lib.fasta_oracle.restype = ctypes.c_int64
        self.lib.fasta_oracle.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 6

    def parse(self, text):
        """(blocks uint16, offset int64, size int32, desc_start int64, desc_end int64)
        of the text (bytes), in the layout of the pairwise sequence store."""
        t = np.frombuffer(bytes(text), dtype=np.uint8) if len(text) else np.zeros(1, np.uint8)
        nb = ctypes.c_int64()
        r = self.lib.fasta_oracle(t.ctypes.data, len(text), None, None, None, None, None, ctypes.byref(nb))
        blocks = np.zeros(max(nb.value, 1), np.uint16)
        arrays = [np.zeros(max(r, 1), dt) for dt in (np.int64, np.int32, np.int64, np.int64)]
        self.lib.fasta_oracle(t.ctypes.data, len(text), blocks.ctypes.data, *[a.ctypes.data for a in arrays],
                              ctypes.byref(nb))
        return (blocks[:nb.value],) + tuple(a[:r] for a in arrays)


class ReferenceFasta:
    """ctypes wrapper of oracle/_ref/libref_fasta.so (oracle/Makefile, target
    `ref`): the reference's own FASTA parser over a file.  Present only where
    it was built."""

    PATH = os.path.join(HERE, "_ref", "libref_fasta.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self):
        self.lib = ctypes.CDLL(self.PATH)
        self.lib.ref_fasta_load.restype = ctypes.c_int64
        self.lib.ref_fasta_load.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 7

    def load(self, path):
        """(blocks, offset, size, descriptions) of the file, or None on error."""
        nb, db = ctypes.c_int64(), ctypes.c_int64()
        name = os.fsencode(path)
        r = self.lib.ref_fasta_load(name, None, None, None, None, None, ctypes.byref(nb), ctypes.byref(db))
        if r < 0:
            return None
        blocks = np.zeros(max(nb.value, 1), np.uint16)
        offset = np.zeros(max(r, 1), np.int64)
        size = np.zeros(max(r, 1), np.int32)
        desc = ctypes.create_string_buffer(max(db.value, 1))
        doff = np.zeros(r + 1, np.int64)
        self.lib.ref_fasta_load(name, blocks.ctypes.data, offset.ctypes.data, size.ctypes.data, desc,
                                doff.ctypes.data, ctypes.byref(nb), ctypes.byref(db))
        descs = [desc.raw[doff[i]:doff[i + 1]] for i in range(r)]
        return blocks[:nb.value], offset[:r], size[:r], descs
