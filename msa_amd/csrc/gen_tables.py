"""Regenerates the scoring-table data files from the reference's table source.

The reference keeps seven 25x25 substitution tables as float literals plus a
name -> (table, gap penalty) registry (src/pairwise/table.cu:41-238 for the
data, :254-262 for the registry).  The tables are data that must stay
identical, quirks included (pam250 has the G and E rows/columns of textbook
PAM250 exchanged, blosum45/blosum50 are permuted), so they are transcribed by
parsing that file rather than retyped.

Outputs (both committed):
  msa_amd/csrc/scoring_tables.inc  int8 data for the C++ host library
  oracle/tables.json               the same data for the oracle and the tests

Usage:  python msa_amd/csrc/gen_tables.py [path/to/src/pairwise/table.cu]
"""
import hashlib
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DEFAULT_SRC = "/root/reference/src/pairwise/table.cu"


def parse_reference_tables(text):
    """Returns [(name, penalty, 25x25 int list)] in registry order."""
    start = text.index("table_data[] = {")
    end = text.index("typedef struct", start)
    body = text[start:end]
    blocks = re.split(r"/\*\s*\[(\d+)\]\s*(\w+)\s*\*/", body)[1:]
    tables = {}
    for k in range(0, len(blocks), 3):
        index, name, chunk = int(blocks[k]), blocks[k + 1], blocks[k + 2]
        rows = re.findall(r"\{([-\d,\s]+)\}", chunk)
        matrix = [[int(v) for v in row.split(",")] for row in rows[:25]]
        assert len(matrix) == 25 and all(len(r) == 25 for r in matrix), name
        tables[index] = (name, matrix)
    registry = re.findall(r'\{"(\w+)",\s*\{&table_data\[(\d+)\],\s*(\d+)\}\}', text)
    out = []
    for name, index, penalty in registry:
        tname, matrix = tables[int(index)]
        assert tname == name, (tname, name)
        out.append((name, int(penalty), matrix))
    return out


def digest(entries):
    h = hashlib.sha256()
    for name, penalty, matrix in entries:
        h.update(name.encode())
        h.update(penalty.to_bytes(1, "little", signed=True))
        h.update(bytes((v & 0xFF) for row in matrix for v in row))
    return h.hexdigest()


def write_outputs(entries):
    lines = [
        "// Generated by msa_amd/csrc/gen_tables.py from the reference's",
        "// src/pairwise/table.cu:41-262.  Do not edit by hand.",
        "// Layout: kTableData[t][row_code][col_code], codes in encoder order",
        "// A C T G R N D Q E H I L K M F P S W Y V B J Z X * (0..24).",
        "// sha256 of (name, penalty, 625 int8) over all tables: " + digest(entries),
        "static const int kTableCount = %d;" % len(entries),
        "static const char* const kTableNames[kTableCount] = {%s};"
        % ", ".join('"%s"' % e[0] for e in entries),
        "static const signed char kTablePenalty[kTableCount] = {%s};"
        % ", ".join(str(e[1]) for e in entries),
        "static const signed char kTableData[kTableCount][25][25] = {",
    ]
    for name, _, matrix in entries:
        lines.append("  { // %s" % name)
        for row in matrix:
            lines.append("    {" + ",".join("%3d" % v for v in row) + "},")
        lines.append("  },")
    lines.append("};")
    with open(os.path.join(HERE, "scoring_tables.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    data = {
        "source": "src/pairwise/table.cu:41-262",
        "sha256": digest(entries),
        "tables": [{"name": n, "penalty": p, "matrix": m} for n, p, m in entries],
    }
    with open(os.path.join(ROOT, "oracle", "tables.json"), "w") as f:
        json.dump(data, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    with open(src) as f:
        entries = parse_reference_tables(f.read())
    write_outputs(entries)
    print("tables:", [e[0] for e in entries], "sha256", digest(entries))
