"""Writes tests/golden/mock_scores.json: the pairwise scores of the reference's
three FASTA fixtures (test/fixtures/mock{1,2,3}.fasta, copied here) under the
tables its own pairwise test uses (test/pairwise.py:13-30: default, blosum62,
pam250) -- 23 pairs x 3 tables = 69 scores.

The reference stores no expected scores (its test compares two
implementations), so each score is produced four independent ways and must
agree before it is written:
  1. oracle/nw_oracle.c, the C restatement of sequential.cpp:29-95;
  2. oracle.museqa_ref.needleman, the restatement of needleman.py:19-63;
  3. oracle.museqa_ref.textbook_nw, textbook global alignment (full matrix);
  4. oracle/_ref/libref_sequential.so, the reference's own sequential.cpp
     compiled from its sources (used when present).

Also writes tests/golden/loader_md5.json with the md5 digests of the decoded
fixture records that the reference's loader test holds (test/loader.py:14-36).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import museqa_ref as M  # noqa: E402

FIXTURES = ["mock1.fasta", "mock2.fasta", "mock3.fasta"]
TABLES = ["default", "blosum62", "pam250"]

# test/loader.py:14-36 of the reference: md5 of str(contents) per record.
LOADER_MD5 = {
    "mock1.fasta": {
        "NC_007728.1": "50e1a4367bd08aa24b96dec226e54b61", "AB492213.1": "40def109d76a2e7ad42b4c6554cf2843",
        "AY186732.1": "1aafda82b3084e002ee47ec466c648fb", "AY590471.1": "b97c3f6adf77956a6111b67380ed2433",
        "NC_032975.1": "7bdc5d496aed747991b11d0d182b0eed"},
    "mock2.fasta": {
        "AF241168.1": "c638e233669d5a37e88eacfad7e3d435", "NC_011560.1": "19cf53404d9129ade12e7d97d82f2b8a",
        "NC_001874.1": "95bd6fb267dcab96ea781e90dcabfbe4"},
    "mock3.fasta": {
        "HQ541801.1": "81c8ea16a4fdf2d13ecb69048a98bd1b", "L03295.1": "6f4603a0d141561f2d73dcbf71b529ea",
        "KY785481.1": "d198d09fdd802b5ddae606b569299fda", "KY325470.1": "aba55fc4c6a7604cc7329bd7b411f00f",
        "NC_032632.1": "e25645acbbcd44f62088e366afa13984"},
}


def main():
    tables = M.load_tables()
    orc = M.CpuOracle()
    ref = M.ReferenceBuild() if M.ReferenceBuild.available() else None
    out = {"generator": "tests/golden/make_golden.py", "reference_build_used": ref is not None, "scores": {}}
    for fx in FIXTURES:
        recs = M.parse_fasta(os.path.join(HERE, fx))
        seqs = [M.encode(s) for _, s in recs]
        blocks = np.array([b for s in seqs for b in s], dtype=np.uint16)
        sizes = np.array([len(s) for s in seqs], dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        pairs = np.array(M.pair_order(len(seqs)), dtype=np.int32)
        for tname in TABLES:
            T, pen = tables[tname]
            c_scores = orc.align(blocks, offsets, sizes, T.astype(np.float32), pen)
            py_scores = np.array(M.needleman(seqs, T.tolist(), pen), dtype=np.float32)
            tb_scores = []
            for i, j in pairs:
                one, two = (i, j) if sizes[i] > sizes[j] else (j, i)
                a = [M.encode_char(ch) for ch in recs[one][1]]
                b = [M.encode_char(ch) for ch in recs[two][1]]
                tb_scores.append(M.textbook_nw(a, b, T, pen))
            tb_scores = np.array(tb_scores, dtype=np.float32)
            assert np.array_equal(c_scores, py_scores), (fx, tname)
            assert np.array_equal(c_scores, tb_scores), (fx, tname)
            if ref is not None:
                assert np.array_equal(c_scores, ref.align_list(blocks, offsets, sizes, pairs, T, pen)), (fx, tname)
            out["scores"]["%s/%s" % (fx, tname)] = {
                "descriptions": [d for d, _ in recs],
                "pairs": pairs.tolist(),
                "scores": [int(v) for v in c_scores],
            }
    with open(os.path.join(HERE, "mock_scores.json"), "w") as f:
        json.dump(out, f, indent=1)
    with open(os.path.join(HERE, "loader_md5.json"), "w") as f:
        json.dump({"source": "test/loader.py:14-36", "digests": LOADER_MD5}, f, indent=1)
    n = sum(len(v["scores"]) for v in out["scores"].values())
    print("wrote %d golden scores (reference build used: %s)" % (n, ref is not None))


if __name__ == "__main__":
    main()
