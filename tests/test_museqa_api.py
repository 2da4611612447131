"""The museqa surface (C++ host layer + Python binding), mirroring the
reference's own tests: test/sequence.py, test/database.py, test/loader.py
and test/pairwise.py.  The pairwise cases need a GPU and compare every
registered algorithm with the oracle, as test/pairwise.py:46-66 compares
'sequential'/'hybrid' with the Python needleman."""
import hashlib
import json
import os

import numpy as np
import pytest

from msa_amd.museqa import database as mdb
from msa_amd.museqa import pairwise as mpw
from msa_amd.museqa import sequence as mseq
from oracle import museqa_ref as M
from tests import helpers as H
from tests.test_oracle import SEQUENCE_FIXTURES

Database, Sequence, encode = mdb.Database, mseq.Sequence, mseq.encode


# ---- test/sequence.py -------------------------------------------------------

@pytest.fixture(params=SEQUENCE_FIXTURES)
def sequence(request):
    return Sequence(request.param), encode(request.param)


def testCanInstantiateSequence(sequence):
    target, expected = sequence
    assert expected == str(target)


def testCanAccessSequenceElement(sequence):
    target, expected = sequence
    for i in range(len(expected)):
        assert expected[i] == target[i]


def testIfSequenceHasCorrectLength(sequence):
    target, expected = sequence
    assert len(expected) == target.length


def testIfRaisesOnIndexError(sequence):
    target, _ = sequence
    with pytest.raises(IndexError):
        target[target.length]


def test_encode_matches_oracle_encoder():
    for text in SEQUENCE_FIXTURES + ["acgt*-x1", "", "NNX\r"]:
        assert encode(text) == M.decode(M.encode(text))
        assert str(Sequence(text)) == M.decode(M.encode(text))


# ---- test/database.py -------------------------------------------------------

DB_FIXTURES = [
    {"SQ01": "MNNQRKKTGRPSFNMLKRARNRVSTGSQLAKRFSKGLLSGQGPMKLVMAFIAFLRFLAIPPTAGILA**",
     "SQ02": "TEVKGYTKGXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXXX**",
     "SQ03": "MKNPKKKSGGFRIVNMLKRGVARVSPFGGLKRLPAGLLLGHGPIRMVLAILAFLRFTAIKPSLGLIN**",
     "SQ04": "AAAAATCGATCAAGCAATCTCTCAGTCTCTCAGTTTACAGCAATCTTTGCAAACACACGAACCACAA**",
     "SQ05": "LLNRFTMAHMKPTYERDVDLGAGTRHVAVEPEVANLDIIGQRIENIKNEHKSTWHYDEDNPYKTWAY**"},
    {"SQ06": "GATAAAAGAACCTATAATCCCTTCGCACACCGCGTCACACCGCGCTATATGCTGCTCATTAGGAAT",
     "SQ07": "GAGCAATCGTTACATAACATGCATTCCAAATAAGGGAATGCGGAAATCCATACTTGGGACTTACAT",
     "SQ08": "GAGCAATCGTTACATAACATGCATTCTAACTAAGGCAATGCGGGATTCCATACTTGGGACTTACAT",
     "SQ09": "GCTCCTTTTTTGTGGATACAATCTCTTGTATACGATATACTTATTGTTAATTTCATTGACCTTTAC",
     "SQ10": "TGCATTCCATATAAGGAATACGGATATGCCCATGTTTGTATCCAAACAGGCGGTCTCCCAGACTCC"},
]


@pytest.fixture(params=range(len(DB_FIXTURES)))
def database(request):
    return DB_FIXTURES[request.param]


@pytest.fixture(params=[2, 3, 4])
def subsets(request, database):
    items = [*database.items()]
    return items[:request.param], items[request.param:]


def testIfCanAddSequences(database):
    db = Database()
    db.add(database)
    for name, seq in database.items():
        assert name == db[name].description
        assert seq == str(db[name].contents)
    assert len(database) == db.count


def testIfCanMergeDatabases(subsets):
    dbs = []
    for i, subset in enumerate(subsets):
        dbs.append(Database())
        dbs[i].add(subset)
        assert len(subset) == dbs[i].count
    for i in range(1, len(dbs)):
        dbs[0].merge(dbs[i])
    for subset in subsets:
        for name, seq in subset:
            assert name == dbs[0][name].description
            assert seq == str(dbs[0][name].contents)
    assert sum(len(s) for s in subsets) == dbs[0].count


def test_only_and_anonymous_entries():
    db = Database()
    db.add(["ACGT", "GGCC"])
    db.add(("named", "TTAA"))
    assert db.count == 3 and db[0].description.startswith("anonymous#")
    sub = db.only([2, 0])
    assert sub.count == 2 and sub[0].description == db[0].description and sub[1].description == "named"
    with pytest.raises(TypeError):
        db.only([0, "named"])
    with pytest.raises(RuntimeError):
        db[7]
    with pytest.raises(RuntimeError):
        db["missing"]


# ---- test/loader.py ---------------------------------------------------------

def testLoadFromFastaFile():
    with open(os.path.join(H.GOLDEN, "loader_md5.json")) as f:
        digests = json.load(f)["digests"]
    for fixture, expected in digests.items():
        db = Database.load(os.path.join(H.GOLDEN, fixture))
        for key, digest in expected.items():
            assert key == db[key].description
            assert digest == hashlib.md5(str(db[key].contents).encode()).hexdigest()


def testIfRaisesOnUnknownFile():
    with pytest.raises(RuntimeError):
        Database.load("unknown/file/path")
    with pytest.raises(RuntimeError):
        Database.load("missing-file.fasta")
    # The GPU loader checks the name and the file before it touches a device.
    with pytest.raises(RuntimeError, match="unknown database parser"):
        Database.load("unknown/file/path", device=True)
    with pytest.raises(RuntimeError, match="does not exist"):
        Database.load("missing-file.fasta", device=True)


# ---- pairwise surface (no kernel launch) -------------------------------------

def test_scoring_table_surface():
    assert mpw.ScoringTable.list() == ["default", "blosum62", "blosum45", "blosum50", "blosum80", "blosum90", "pam250"]
    t = mpw.ScoringTable("pam250")
    assert t.penalty == 8 and t["W", "W"] == 17 and t[3, 3] == 4
    with pytest.raises(RuntimeError, match="unknown pairwise scoring table"):
        mpw.ScoringTable("nope")
    with pytest.raises(RuntimeError):
        t[25, 0]


def test_algorithm_registry():
    names = mpw.list()
    for n in ["default", "needleman", "hybrid", "needleman-hybrid", "sequential", "distributed",
              "needleman-sequential", "needleman-distributed", "hip", "hip-plain", "hip-batched", "hip-long"]:
        assert n in names
    with pytest.raises(RuntimeError, match="unknown pairwise algorithm 'nope'"):
        mpw.run(Database(), algorithm="nope")


def test_distance_matrix_layout():
    """[x, y] reads buf[nchoose(max) + min]; the diagonal reads 0 (pairwise.cuh:99-107)."""
    scores = [float(v) for v in range(10)]
    d = mpw.DistanceMatrix(scores, 5)
    assert d.count == 5
    c = 0
    for i in range(1, 5):
        for j in range(i):
            assert d[i, j] == d[j, i] == scores[c]
            c += 1
    assert all(d[i, i] == 0 for i in range(5))


def test_python_callable_algorithm_is_accepted():
    """pairwise.pyx:97-98: a Python function may stand for the algorithm."""
    db = Database.load(os.path.join(H.GOLDEN, "mock2.fasta"))

    def fake(db, table):
        return mpw.DistanceMatrix([1.0] * 3, db.count)

    assert mpw.run(db, algorithm=fake)[1, 0] == 1.0


# ---- test/pairwise.py (GPU) ---------------------------------------------------

def python_needleman(db, table):
    """The reference's algorithm.needleman restated (oracle), as a callable."""
    seqs = [db[i].contents.native.blocks() for i in range(db.count)]
    T = [[table[x, y] for y in range(25)] for x in range(25)]
    return mpw.DistanceMatrix(M.needleman(seqs, T, table.penalty), db.count)


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["mock1.fasta", "mock2.fasta", "mock3.fasta"])
@pytest.mark.parametrize("table", ["default", "blosum62", "pam250"])
@pytest.mark.parametrize("algorithm", ["default", "hybrid", "sequential", "hip-batched", "hip-plain"])
def testNeedlemanAlgorithms(fixture, table, algorithm):
    db = Database.load(os.path.join(H.GOLDEN, fixture))
    expected = mpw.run(db, algorithm=python_needleman, table=table)
    produced = mpw.run(db, algorithm=algorithm, table=table)
    assert expected.count == produced.count
    for i in range(expected.count):
        for j in range(produced.count):
            assert expected[i, j] == produced[i, j]
    golden = H.golden_scores()["%s/%s" % (fixture, table)]["scores"]
    assert [int(v) for v in produced.scores()] == golden


@pytest.mark.gpu
def test_module_pairwise_contract():
    """module::pairwise::check/run (src/pairwise.cpp:27-55)."""
    from msa_amd.museqa._core import core
    db = Database.load(os.path.join(H.GOLDEN, "mock3.fasta"))
    opts = {"pairwise": "needleman", "scoring-table": "blosum62"}
    assert core.module_check(opts)
    with pytest.raises(RuntimeError, match="unknown pairwise algorithm chosen"):
        core.module_check({"pairwise": "zzz"})
    out_db, dm, count = core.module_run(opts, db.native)
    assert count == 5 and dm.count == 5
    assert [int(v) for v in dm.raw()] == H.golden_scores()["mock3.fasta/blosum62"]["scores"]


# ---- phylogeny surface ---------------------------------------------------------

PHYLOGENY_NAMES = {
    # src/phylogeny/phylogeny.cpp:26-41: name -> matrix-removal variant
    "default": 0, "njoining": 0, "hybrid": 0, "linear": 1, "symmetric": 0, "njoining-hybrid": 0,
    "njoining-linear": 1, "njoining-symmetric": 0, "sequential": 0, "distributed": 0,
    "njoining-sequential": 0, "njoining-sequential-linear": 1, "njoining-distributed": 0,
    "njoining-distributed-linear": 1, "hip": 0, "hip-linear": 1,
}


def test_phylogeny_registry():
    from msa_amd.museqa import phylogeny as mpg
    from msa_amd.museqa._core import core
    assert sorted(mpg.list()) == sorted(PHYLOGENY_NAMES)
    with pytest.raises(RuntimeError, match="unknown phylogeny algorithm 'nope'"):
        mpg.run(mpw.DistanceMatrix([1.0], 2), algorithm="nope")
    assert core.phylogeny_module_check({"phylogeny": "njoining-linear"})
    with pytest.raises(RuntimeError, match="unknown phylogeny algorithm chosen: 'zzz'"):
        core.phylogeny_module_check({"phylogeny": "zzz"})


def test_phylogeny_trivial_counts():
    """No join happens below two sequences: no tree at 0, a lone leaf at 1."""
    from msa_amd.museqa import phylogeny as mpg
    d = mpw.DistanceMatrix([4.0, 5.0, 6.0], 3)
    assert len(mpg.run(d, count=0)) == 0
    one = mpg.run(d, count=1)
    assert len(one) == 1 and one.leaves == 1 and one.root == 0
    assert one[0]["parent"] == mpg.UNDEFINED and one[0]["child"] == (mpg.UNDEFINED, mpg.UNDEFINED)
    with pytest.raises(RuntimeError, match="fewer sequences"):
        mpg.run(d, count=4)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(PHYLOGENY_NAMES))
def test_phylogeny_algorithms_match_oracle(name):
    from msa_amd.museqa import phylogeny as mpg
    rng = np.random.default_rng(11)
    n = 57
    scores = rng.integers(-500, 500, n * (n - 1) // 2).astype(np.float32)
    tree = mpg.run(mpw.DistanceMatrix(scores.tolist(), n), algorithm=name)
    want = M.NjOracle().build(scores, n, PHYLOGENY_NAMES[name])
    assert tree.leaves == n and tree.root == 2 * n - 2
    H.assert_same_tree(tree.nodes(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["mock1.fasta", "mock2.fasta", "mock3.fasta"])
def test_module_pipeline_pairwise_then_phylogeny(fixture):
    """module::pairwise's conduit into module::phylogeny (src/phylogeny.cpp:25-34):
    the tree is the oracle's over the golden scores."""
    from msa_amd.museqa import phylogeny as mpg
    from msa_amd.museqa._core import core
    db = Database.load(os.path.join(H.GOLDEN, fixture))
    out_db, dm, count = core.module_run({"scoring-table": "blosum62"}, db.native)
    for name, variant in [("default", 0), ("njoining-linear", 1)]:
        tdb, tree = core.phylogeny_module_run({"phylogeny": name}, out_db, dm, count)
        assert tdb.count == db.count and tree.count == max(0, 2 * count - 1)
        golden = np.array(H.golden_scores()["%s/blosum62" % fixture]["scores"], dtype=np.float32)
        want = M.NjOracle().build(golden, count, variant)
        H.assert_same_tree(mpg.GuideTree(tree).nodes(), want)


# ---- museqa.pgalign (src/pgalign/pgalign.cpp:24-66, src/pgalign.cpp:25-47) -------

def test_pgalign_registry():
    from msa_amd.museqa import pgalign as mpa
    from msa_amd.museqa import phylogeny as mpg
    from msa_amd.museqa._core import core
    assert sorted(mpa.list()) == sorted(["default", "myers", "sequential", "myers-sequential", "hip"])
    tree = mpg.run(mpw.DistanceMatrix([4.0], 2), count=1)
    db = Database()
    db.add(("a", "ACGT"))
    with pytest.raises(RuntimeError, match="unknown pgalign algorithm 'nope'"):
        mpa.run(db, tree, algorithm="nope")
    assert core.pgalign_module_check({"pgalign": "myers"})
    with pytest.raises(RuntimeError, match="unknown pgalign algorithm chosen: 'zzz'"):
        core.pgalign_module_check({"pgalign": "zzz"})


def test_pgalign_single_sequence_needs_no_join():
    """One sequence is its own alignment (no GPU needed): residue i in column
    i, padding excluded as unpadded() counts it."""
    from msa_amd.museqa import pgalign as mpa
    from msa_amd.museqa import phylogeny as mpg
    db = Database()
    db.add(("a", "MKVLA"))
    tree = mpg.run(mpw.DistanceMatrix([], 1), count=1)
    al = mpa.run(db, tree)
    assert len(al) == 1 and al.length == 5 and al[0] == "MKVLA" and al.columns(0) == [0, 1, 2, 3, 4]
    assert al.at(0, 2) == "V" and al.at(0, 7) == "*"


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["mock1.fasta", "mock2.fasta", "mock3.fasta"])
def test_module_pipeline_pairwise_phylogeny_pgalign(fixture):
    """module::pairwise -> module::phylogeny -> module::pgalign: every name
    gives the oracle's alignment along the tree, rows read back with gaps."""
    from msa_amd.museqa import pgalign as mpa
    from msa_amd.museqa import phylogeny as mpg
    from msa_amd.museqa._core import core
    db = Database.load(os.path.join(H.GOLDEN, fixture))
    opts = {"scoring-table": "blosum62"}
    out_db, dm, count = core.module_run(opts, db.native)
    tdb, tree = core.phylogeny_module_run(opts, out_db, dm, count)
    nodes = mpg.GuideTree(tree).nodes()
    blocks, offsets, sizes = H.fixture_db(fixture)[1]
    T, p = H.tables()["blosum62"]
    cols, length, _ = M.PgOracle().align(blocks, offsets, sizes, nodes, T, p)
    for name in mpa.list():
        al = mpa.Alignment(core.pgalign_module_run(dict(opts, pgalign=name), tdb, tree))
        assert len(al) == count and al.length == length
        got = np.concatenate([np.array(al.columns(i), dtype=np.int32) for i in range(count)])
        assert np.array_equal(got, cols)
        residues = M.PgOracle().unpadded(blocks, offsets, sizes)
        for i in range(count):
            row = al[i]
            assert len(row) == length
            assert row.replace("-", "") == str(db[i].contents)[:int(residues[i])]
