"""Writes tests/golden/mc_cases.json from the reference's marching-cubes tables.

Runs only in a checkout that has the reference sources (they are read as text):
    python tests/golden/make_mc_fixture.py [/path/to/reference]

The fixture is data: for each HPMC corner code (bit x+2y+4z set when the
corner sample is below the iso-value) the vertex count and, per vertex slot,
the cell edge as (shift_x, shift_y, shift_z, axis).  It restates
  - hpmc/src/tables.cpp:38-296  (classic triangle table, classic case index)
  - hpmc/src/tables.cpp:299-307 (classic edge -> shift and axis)
  - hpmc/src/constants.cpp:42-50 (classic case index -> HPMC code)
  - hpmc/src/constants.cpp:255-264, 304-326 (tables indexed by HPMC code)
and a SHA-256 over a canonical text form so later checks can pin it.
"""
import hashlib
import json
import os
import re
import sys


def _classic_to_hpmc(code):
    # Classic corner j of the triangle table sits at HPMC bit position:
    #   v0=(0,0,0)->0, v1=(1,0,0)->1, v2=(1,0,1)->5, v3=(0,0,1)->4,
    #   v4=(0,1,0)->2, v5=(1,1,0)->3, v6=(1,1,1)->7, v7=(0,1,1)->6
    dest = [0, 1, 5, 4, 2, 3, 7, 6]
    out = 0
    for j in range(8):
        if (code >> j) & 1:
            out |= 1 << dest[j]
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "hpmc/src/tables.cpp")).read()
    tri_block = text[text.index("HPMC_triangle_table"):text.index("HPMC_edge_table")]
    rows = re.findall(r"\{([-\d,\s]+)\}", tri_block)
    tri = [[int(v) for v in r.split(",")] for r in rows]
    assert len(tri) == 256 and all(len(r) == 16 for r in tri), "triangle table shape"
    edge_block = text[text.index("HPMC_edge_table"):text.index("HPMC_midpoint_table")]
    erows = re.findall(r"\{\s*([-\d.]+)\s*,\s*([-\d.]+)\s*,\s*([-\d.]+)\s*,\s*([-\d.]+)f?\s*\}", edge_block)
    edges = [[int(float(v)) for v in r] for r in erows]
    assert len(edges) == 12, "edge table shape"

    cases = [None] * 256
    for classic in range(256):
        row = tri[classic]
        n = 0
        while n < 16 and row[n] != -1:
            n += 1
        slots = [edges[row[i]] for i in range(n)]
        cases[_classic_to_hpmc(classic)] = {"count": n, "classic_edges": row[:n], "slots": slots}
    assert all(c is not None for c in cases)

    canon = "\n".join(
        "%d:%d:%s" % (code, c["count"], ";".join(",".join(str(v) for v in s) for s in c["slots"]))
        for code, c in enumerate(cases))
    doc = {
        "description": "Marching-cubes cases indexed by HPMC corner code (bit x+2y+4z set when f<T); "
                       "slots are (shift_x, shift_y, shift_z, axis) with axis 0=x,1=y,2=z. "
                       "Restated from hpmc/src/tables.cpp:38-307 and constants.cpp:42-50,255-326.",
        "classic_edge_table": edges,
        "sha256": hashlib.sha256(canon.encode()).hexdigest(),
        "cases": cases,
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mc_cases.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=0, separators=(",", ":"))
    print("wrote", out, doc["sha256"])


if __name__ == "__main__":
    main()
