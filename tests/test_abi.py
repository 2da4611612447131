"""The C-ABI library (CPU-side checks: load, exports, argument errors).

Nothing here launches a kernel; the calls exercised fail or finish on the
host before any device work.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from msa_amd import _native as N
from msa_amd import workloads as W
from oracle import museqa_ref as M
from tests import helpers as H

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "msa_hip.h")


def declared_symbols():
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, name
    assert set(N.SIGNATURES) == set(names)


def test_graph_entry_points_check_arguments_first():
    """msa_graph_*: the null stream, a null graph and a null output are
    refused before any device call; destroying no graph is a no-op."""
    lib = N.lib()
    g = ctypes.c_void_p()
    assert lib.msa_graph_capture_begin(None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_graph_capture_end(None, ctypes.byref(g)) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_graph_capture_end(ctypes.c_void_p(1), None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_graph_launch(None, None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_graph_destroy(None) == 0


def test_error_strings():
    lib = N.lib()
    for code in range(0, 11):
        assert lib.msa_error_string(code)
    assert lib.msa_error_string(0) == b"no error"


def test_table_registry_matches_oracle_tables():
    """scoring_table::list/make (table.cu:254-330)."""
    assert N.table_names() == ["default", "blosum62", "blosum45", "blosum50", "blosum80", "blosum90", "pam250"]
    for name, (T, p) in M.load_tables().items():
        t, pen = N.table(name)
        assert pen == p and (t == T).all(), name


def test_unknown_table_is_an_error():
    with pytest.raises(N.MsaError) as e:
        N.table("blosum99")
    assert e.value.status == N.MSA_ERR_NOT_FOUND
    assert "unknown pairwise scoring table 'blosum99'" in str(e.value)


def test_gap_code_is_refused_before_any_device_work():
    db = H.strings_db(["ACGT-A", "ACGT"])
    T, p = H.tables()["blosum62"]
    with pytest.raises(N.MsaError) as e:
        N.align(*db, T, p)
    assert e.value.status == N.MSA_ERR_GAP_CODE


def test_table_outside_byte_profile_is_refused():
    db = H.strings_db(["ACGT", "ACGT"])
    T = np.full((25, 25), -20, dtype=np.int32)
    with pytest.raises(N.MsaError) as e:
        N.align(*db, T, 4)
    assert e.value.status == N.MSA_ERR_TABLE


def test_all_pairs_count_is_checked():
    lib = N.lib()
    db = H.strings_db(["ACGT", "ACGT", "AC"])
    blocks, offsets, sizes = db
    T, p = H.tables()["blosum62"]
    out = np.zeros(8, dtype=np.int32)
    opts = N.options()
    rc = lib.msa_nw_align(blocks.ctypes.data, offsets.ctypes.data, sizes.ctypes.data, 3, None, 5,
                          T.ctypes.data, p, ctypes.byref(opts), out.ctypes.data)
    assert rc == N.MSA_ERR_INVALID_ARGUMENT
    assert b"nseq*(nseq-1)/2" in lib.msa_last_error()


def test_null_arguments():
    lib = N.lib()
    assert lib.msa_nw_plan_create(None, None, None, None, 0, None, 0, None, 0, None, None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_device_count(None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_nw_plan_destroy(None) == N.MSA_OK


def test_device_count_is_callable():
    n = N.device_count()
    assert n >= 0


def _tuning_status(tmp_path, text):
    path = tmp_path / "t.json"
    path.write_text(text)
    blocks, offsets, sizes = W.pack([np.array([0, 1, 2], dtype=np.uint8), np.array([3, 1], dtype=np.uint8)])
    table = np.zeros(625, dtype=np.int32)
    out = np.zeros(1, dtype=np.int32)
    opts = N.options(tuning_file=str(path))
    rc = N.lib().msa_nw_align(blocks.ctypes.data, offsets.ctypes.data, sizes.ctypes.data, 2, None, 1,
                              table.ctypes.data, 4, ctypes.byref(opts), out.ctypes.data)
    return rc, N.lib().msa_last_error().decode()


@pytest.mark.parametrize("text,needle", [
    ("not json", "not a JSON object"),
    ('{"format": "other/1", "long_threshold": 8192, "bins": [{"min_query": 0, "lanes": 16, "rows": 32}]}',
     "format"),
    ('{"format": "msa-nw-tuning/1", "bins": [{"min_query": 0, "lanes": 16, "rows": 32}]}', "long_threshold"),
    ('{"format": "msa-nw-tuning/1", "long_threshold": 8192, "bins": []}', "non-empty"),
    ('{"format": "msa-nw-tuning/1", "long_threshold": 8192, "bins": [{"min_query": 0, "lanes": 64, "rows": 32}]}',
     "does not have"),
    ('{"format": "msa-nw-tuning/1", "long_threshold": 8192, "bins": [{"min_query": 5, "lanes": 16, "rows": 32}]}',
     "start at 0"),
    ('{"format": "msa-nw-tuning/1", "long_threshold": 8192, "bins": [{"min_query": 0, "lanes": 16, "rows": 32},'
     ' {"min_query": 0, "lanes": 16, "rows": 16}]}', "increase"),
])
def test_tuning_table_errors_are_reported_before_any_device_call(tmp_path, text, needle):
    rc, msg = _tuning_status(tmp_path, text)
    assert rc == N.MSA_ERR_INVALID_ARGUMENT
    assert needle in msg


def test_valid_tuning_table_is_accepted(tmp_path):
    """A well-formed table passes the reader: what fails next (here, with no
    GPU) is the device, not the table."""
    rc, msg = _tuning_status(tmp_path, '{"format": "msa-nw-tuning/1", "device": "x", "long_threshold": 4096, '
                                       '"bins": [{"min_query": 0, "lanes": 16, "rows": 32}, '
                                       '{"min_query": 256, "lanes": 32, "rows": 16}], "measurements": []}')
    assert rc != N.MSA_ERR_INVALID_ARGUMENT, msg


def test_nj_argument_errors_before_any_device_call():
    lib = N.lib()
    nodes = np.zeros(1, dtype=N.NJ_NODE)
    assert lib.msa_nj_build(None, 70000, 0, None) == N.MSA_ERR_TOO_LARGE
    assert lib.msa_nj_build(None, 5, 7, None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_nj_build(None, -1, 0, None) == N.MSA_ERR_INVALID_ARGUMENT
    assert lib.msa_nj_build_device(None, 5, 2, None, None) == N.MSA_ERR_INVALID_ARGUMENT
    # Fewer than two OTUs: no tree, nothing to do.
    assert lib.msa_nj_build(None, 1, 0, nodes.ctypes.data) == N.MSA_OK
    assert lib.msa_nj_build(None, 0, 1, None) == N.MSA_OK
