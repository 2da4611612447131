"""Deterministic synthetic workloads for the all-pairs Needleman-Wunsch stage.

Every draw comes from SplitMix64 (Steele, Lea & Flood 2014) with explicit
transforms, so a (workload, seed) pair names one input on every machine:
  uniform(0,1)   = (x >> 11) * 2^-53
  integer in [a,b] = a + floor(uniform * (b - a + 1))
  normal         = Box-Muller on two uniforms
  categorical    = first index whose cumulative weight exceeds a uniform

Residues are the reference's 5-bit codes (src/encoder.cpp:31-43): A0 C1 T2 G3
R4 N5 D6 Q7 E8 H9 I10 L11 K12 M13 F14 P15 S16 W17 Y18 V19 ...
Sequences are packed three codes per uint16 block at bit shifts 1, 6 and 11,
with bit 0 set on the block that holds padding (src/encoder.cpp:76-98).
"""
import os

import numpy as np

MASK64 = (1 << 64) - 1

# Letter -> code, src/encoder.cpp:31-34 (indexed by letter - 'A').
_ENCODE = (0x00, 0x14, 0x01, 0x06, 0x08, 0x0E, 0x03, 0x09, 0x0A, 0x15, 0x0C, 0x0B, 0x0D,
           0x05, 0x17, 0x0F, 0x07, 0x04, 0x10, 0x02, 0x17, 0x13, 0x11, 0x17, 0x12, 0x16)
END_CODE = 0x18


def letter_code(ch):
    return _ENCODE[ord(ch.upper()) - ord("A")]


AMINO20 = "ARNDCQEGHILKMFPSTWYV"
# BLOSUM62 background frequencies (Henikoff & Henikoff 1992), in AMINO20 order.
BLOSUM62_BACKGROUND = (0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
                       0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073)
AMINO20_CODES = np.array([letter_code(c) for c in AMINO20], dtype=np.uint8)
DNA_CODES = np.array([letter_code(c) for c in "ACGT"], dtype=np.uint8)


class SplitMix64:
    """SplitMix64: state += 0x9E3779B97F4A7C15, then the two xor-multiply rounds."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self, n):
        """n consecutive outputs as a uint64 array."""
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & MASK64
        return z

    def uniform(self, n):
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def integers(self, lo, hi, n):
        """n integers uniform in [lo, hi]."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    def categorical(self, weights, n):
        cdf = np.cumsum(np.asarray(weights, dtype=np.float64))
        cdf /= cdf[-1]
        return np.minimum(np.searchsorted(cdf, self.uniform(n), side="right"), len(cdf) - 1)


def pack(codes_list):
    """Packs code arrays into (blocks uint16, offsets int64, sizes int32)."""
    sizes = np.array([(len(c) + 2) // 3 for c in codes_list], dtype=np.int32)
    offsets = np.zeros(len(codes_list), dtype=np.int64)
    if len(codes_list):
        offsets[1:] = np.cumsum(sizes, dtype=np.int64)[:-1]
    blocks = np.zeros(int(sizes.sum()), dtype=np.uint16)
    for k, codes in enumerate(codes_list):
        n = len(codes)
        nb = int(sizes[k])
        if nb == 0:
            continue
        padded = np.full(3 * nb, END_CODE, dtype=np.uint16)
        padded[:n] = np.asarray(codes, dtype=np.uint16)
        tri = padded.reshape(nb, 3)
        blk = (tri[:, 0] << 1) | (tri[:, 1] << 6) | (tri[:, 2] << 11)
        if n % 3:
            blk[-1] |= 1
        blocks[offsets[k]:offsets[k] + nb] = blk
    return blocks, offsets, sizes


def unpack(blocks, offset, size):
    """Codes of one packed sequence, padding included."""
    b = np.asarray(blocks[offset:offset + size], dtype=np.uint32)
    return np.stack([(b >> 1) & 31, (b >> 6) & 31, (b >> 11) & 31], axis=1).reshape(-1).astype(np.uint8)


def pair_of(c):
    """(i, j) of pair c in the all-pairs order i = 1..n-1, j < i, c = i(i-1)/2 + j
    (src/pairwise/pairwise.cpp:79-81)."""
    i = int((1 + np.sqrt(1 + 8 * float(c))) // 2)
    while i * (i - 1) // 2 > c:
        i -= 1
    while (i + 1) * i // 2 <= c:
        i += 1
    return i, c - i * (i - 1) // 2


class Workload:
    def __init__(self, name, codes, table, penalty, description):
        self.name = name
        self.codes = codes
        self.table = np.asarray(table, dtype=np.int32).reshape(25, 25)
        self.penalty = int(penalty)
        self.description = description
        self.blocks, self.offsets, self.sizes = pack(codes)
        self.lengths = np.array([len(c) for c in codes], dtype=np.int64)

    @property
    def nseq(self):
        return len(self.codes)

    @property
    def npairs(self):
        return self.nseq * (self.nseq - 1) // 2

    def cells(self):
        """Sum over all pairs of L1 * L2 (unpadded)."""
        s = self.lengths.sum()
        return int((s * s - (self.lengths ** 2).sum()) // 2)

    def checksum(self):
        import hashlib
        h = hashlib.sha256(self.blocks.tobytes())
        h.update(self.sizes.tobytes())
        return h.hexdigest()[:16]


def _protein(rng, lengths, background):
    weights = BLOSUM62_BACKGROUND if background else [1.0] * 20
    return [AMINO20_CODES[rng.categorical(weights, int(L))] for L in lengths]


def dna_table(match=5, mismatch=-4):
    """+match/-mismatch over A, C, T, G (codes 0..3); -mismatch elsewhere, 0 at [24][24]."""
    t = np.full((25, 25), mismatch, dtype=np.int32)
    for a in range(4):
        t[a, a] = match
    t[24, 24] = 0
    return t


def uniform_small(tables, seed=1):
    """16 protein sequences, 20 amino acids uniform, L ~ U[100,300], blosum62."""
    rng = SplitMix64(seed)
    lengths = rng.integers(100, 300, 16)
    t, p = tables["blosum62"]
    return Workload("uniform-16", _protein(rng, lengths, False), t, p,
                    "16 protein seqs, uniform residues, L~U[100,300], blosum62 p=4, seed %d" % seed)


def uniform_batch(tables, seed=2, count=512):
    """512 protein sequences, BLOSUM62 background, L ~ U[800,1200]."""
    rng = SplitMix64(seed)
    lengths = rng.integers(800, 1200, count)
    t, p = tables["blosum62"]
    return Workload("uniform-%d" % count, _protein(rng, lengths, True), t, p,
                    "%d protein seqs, BLOSUM62 background, L~U[800,1200], blosum62 p=4, seed %d" % (count, seed))


def lognormal_lengths(rng, count, median=400.0, sigma=1.0, lo=50, hi=5000):
    return np.clip(np.round(np.exp(np.log(median) + sigma * rng.normal(count))), lo, hi).astype(np.int64)


def uneven_batch(tables, seed=3, count=2000):
    """2000 protein sequences, BLOSUM62 background, L = clip(round(lognormal(ln 400, 1)), 50, 5000)."""
    rng = SplitMix64(seed)
    lengths = lognormal_lengths(rng, count)
    t, p = tables["blosum62"]
    return Workload("uneven-%d" % count, _protein(rng, lengths, True), t, p,
                    "%d protein seqs, BLOSUM62 background, L=clip(round(lognormal(ln400,1)),50,5000), "
                    "blosum62 p=4, seed %d" % (count, seed))


def long_dna(seed=4, count=64, length=30000, mutation=0.10, penalty=8):
    """64 DNA sequences: a uniform ancestor plus independent 10% point mutations."""
    rng = SplitMix64(seed)
    ancestor = rng.integers(0, 3, length)
    codes = []
    for _ in range(count):
        s = ancestor.copy()
        hit = rng.uniform(length) < mutation
        shift = rng.integers(1, 3, length)
        s = np.where(hit, (s + shift) % 4, s)
        codes.append(DNA_CODES[s])
    return Workload("dna-%dx%d" % (count, length), codes, dna_table(), penalty,
                    "%d DNA seqs of %d (ancestor + %.0f%% point mutations), +5/-4, gap %d, seed %d"
                    % (count, length, 100 * mutation, penalty, seed))


def serving_jobs(tables, count, seed=5, min_seqs=8, max_seqs=256, lo=100, hi=2000):
    """Jobs of U{8..256} protein sequences with L ~ U[100,2000]."""
    rng = SplitMix64(seed)
    t, p = tables["blosum62"]
    jobs = []
    for j in range(count):
        n = int(rng.integers(min_seqs, max_seqs, 1)[0])
        lengths = rng.integers(lo, hi, n)
        jobs.append(Workload("job-%d" % j, _protein(rng, lengths, True), t, p, "serving job %d" % j))
    return jobs


def poisson_arrivals(rate, count, seed=50):
    """Arrival times (seconds) of a Poisson stream of the given rate."""
    rng = SplitMix64(seed)
    return np.cumsum(-np.log1p(-rng.uniform(count)) / rate)


def fasta_block(rng, target_bytes, width=60, median=400.0, sigma=1.0, lo=50, hi=5000):
    """FASTA text of about `target_bytes`: protein records with lengths
    clip(round(lognormal(ln median, sigma)), lo, hi), BLOSUM62 background
    residues as upper-case letters, `width` residues per line, headers
    '>s' + 12-digit record number.  Built with array arithmetic (no per-record
    Python loop); returns a uint8 array."""
    hdr = 15
    est = int(target_bytes // (hdr + median)) + 64      # more records than needed; cut below
    lengths = lognormal_lengths(rng, est, median, sigma, lo, hi)
    per = hdr + lengths + (lengths + width - 1) // width
    cum = np.cumsum(per)
    count = max(1, int(np.searchsorted(cum, target_bytes, side="right")))
    lengths, per, cum = lengths[:count], per[:count], cum[:count]
    total = int(cum[-1])
    base = cum - per
    out = np.full(total, ord("\n"), dtype=np.uint8)
    out[base] = ord(">")
    out[base + 1] = ord("s")
    ids = np.arange(count, dtype=np.int64)
    for d in range(12):
        out[base + 2 + d] = (ids // 10 ** (11 - d)) % 10 + ord("0")
    n = int(lengths.sum())
    letters = np.frombuffer(AMINO20.encode(), dtype=np.uint8)[rng.categorical(BLOSUM62_BACKGROUND, n)]
    rid = np.repeat(ids, lengths)
    k = np.arange(n, dtype=np.int64) - np.repeat(np.cumsum(lengths) - lengths, lengths)
    out[base[rid] + hdr + k + k // width] = letters
    return out


def fasta_database(seed=6, target_bytes=512 << 20, block_bytes=64 << 20, width=60):
    """A protein FASTA database of about `target_bytes`: one `fasta_block` of
    `block_bytes` (seed `seed`) repeated.  Returns (uint8 text, description)."""
    rng = SplitMix64(seed)
    block = fasta_block(rng, min(block_bytes, target_bytes), width=width)
    reps = max(1, int(round(target_bytes / len(block))))
    text = np.tile(block, reps)
    desc = ("protein FASTA, %d x a %.1f MiB block of records with L=clip(round(lognormal(ln400,1)),50,5000), "
            "BLOSUM62 background, %d residues per line, seed %d" % (reps, len(block) / 2**20, width, seed))
    return text, desc


# ---- the reference's own benchmark databases -----------------------------------

_DB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden", "db")

REFERENCE_CASES = {
    1: "50 complete viral genomes of about 10,000 nt",
    2: "100 complete viral genomes of about 10,000 nt",
    3: "200 complete viral genomes of about 5,000 nt",
    4: "400 complete viral genomes of about 5,000 nt",
    5: "400 complete viral genomes of about 2,000 nt",
    6: "800 complete viral genomes of about 2,000 nt",
}


def reference_case_text(k):
    """The text of the reference's db/case<k>.fasta (kept xz-compressed under
    tests/golden/db/, checked against the original's sha256)."""
    import hashlib
    import json
    import lzma
    name = "case%d.fasta" % k
    with open(os.path.join(_DB, name + ".xz"), "rb") as f:
        text = lzma.decompress(f.read())
    with open(os.path.join(_DB, "digests.json")) as f:
        want = json.load(f)[name]["sha256"]
    if hashlib.sha256(text).hexdigest() != want:
        raise RuntimeError("%s does not match the reference's file" % name)
    return text


class StoredWorkload(Workload):
    """A workload whose sequences come as a packed store (blocks, offsets and
    sizes), e.g. from the FASTA loader; lengths are the residues before the
    padding of the last block."""

    def __init__(self, name, blocks, offsets, sizes, table, penalty, description):
        self.name = name
        self.table = np.asarray(table, dtype=np.int32).reshape(25, 25)
        self.penalty = int(penalty)
        self.description = description
        self.blocks = np.ascontiguousarray(blocks, dtype=np.uint16)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        lengths = []
        for o, s in zip(self.offsets, self.sizes):
            codes = unpack(self.blocks, int(o), int(s))
            n = len(codes)
            while n > len(codes) - 2 and n > 0 and codes[n - 1] == 24:
                n -= 1
            lengths.append(n)
        self.lengths = np.array(lengths, dtype=np.int64)
        self.codes = None

    @property
    def nseq(self):
        return len(self.sizes)


def reference_case(k, tables, table="default"):
    """db/case<k>.fasta parsed on the GPU (msa_fasta_parse), scored with a
    named table (the reference's command line defaults to "default")."""
    from . import _native as N
    text = reference_case_text(k)
    loader = N.FastaLoader()
    try:
        loader.parse(text)
        blocks, offsets, sizes, _, _ = loader.store()
    finally:
        loader.close()
    t, p = tables[table]
    return StoredWorkload("case%d" % k, blocks, offsets, sizes, t, p,
                          "reference db/case%d.fasta: %s, %s p=%d" % (k, REFERENCE_CASES[k], table, p))
