"""Copies the reference's benchmark databases (db/case1..6.fasta: 50-800
complete viral genomes of about 2k, 5k and 10k nucleotides, fetched from
NCBI by the reference's db/ncbi.pl) into tests/golden/db/ as xz-compressed
data, with the sha256 of every original.  The files are data, not source:
the GPU parity tests and bench.py parse them with the GPU FASTA loader.

Run from the repository root where the reference tree exists:
    python tests/golden/make_db_fixtures.py /root/reference/db
"""
import hashlib
import json
import lzma
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    out = os.path.join(HERE, "db")
    os.makedirs(out, exist_ok=True)
    digests = {}
    for k in range(1, 7):
        name = "case%d.fasta" % k
        with open(os.path.join(src, name), "rb") as f:
            data = f.read()
        digests[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
        with open(os.path.join(out, name + ".xz"), "wb") as f:
            f.write(lzma.compress(data, preset=9))
    with open(os.path.join(out, "digests.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/db")
