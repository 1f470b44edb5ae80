"""Marching-cubes case table: fixture integrity and topology invariants.

The fixture (tests/golden/mc_cases.json) restates hpmc/src/tables.cpp:38-307
re-indexed by HPMC code (constants.cpp:42-50).  These checks pin it without
trusting it: counts and edge sets follow from the corner signs, and every
triangle's winding agrees with one orientation convention, so a dropped,
swapped or mis-indexed slot fails here.
"""
import hashlib
import json
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "mc_cases.json")


def _doc():
    return json.load(open(FIXTURE))


def _corner(bit):
    return np.array([bit & 1, (bit >> 1) & 1, (bit >> 2) & 1], float)


def _edge_endpoints(slot):
    sx, sy, sz, axis = slot
    a = np.array([sx, sy, sz], float)
    b = a.copy()
    b[axis] += 1
    return a, b


def test_fixture_hash_and_histogram():
    doc = _doc()
    canon = "\n".join(
        "%d:%d:%s" % (code, c["count"], ";".join(",".join(str(v) for v in s) for s in c["slots"]))
        for code, c in enumerate(doc["cases"]))
    assert hashlib.sha256(canon.encode()).hexdigest() == doc["sha256"]
    counts = [c["count"] for c in doc["cases"]]
    hist = {k: counts.count(k) for k in sorted(set(counts))}
    assert hist == {0: 2, 3: 16, 6: 50, 9: 80, 12: 76, 15: 32}
    assert counts[0] == 0 and counts[255] == 0


def test_edges_are_exactly_the_sign_changes():
    for code, c in enumerate(_doc()["cases"]):
        inside = lambda p: (code >> (int(p[0]) + 2 * int(p[1]) + 4 * int(p[2]))) & 1
        used = set()
        for s in c["slots"]:
            a, b = _edge_endpoints(s)
            assert inside(a) != inside(b), (code, s)
            used.add(tuple(s))
        crossed = set()
        for axis in range(3):
            for u in range(2):
                for v in range(2):
                    sh = [0, 0, 0]
                    others = [q for q in range(3) if q != axis]
                    sh[others[0]], sh[others[1]] = u, v
                    a = np.array(sh, float)
                    b = a.copy()
                    b[axis] += 1
                    if inside(a) != inside(b):
                        crossed.add((sh[0], sh[1], sh[2], axis))
        assert used == crossed, code
        assert c["count"] % 3 == 0


def test_triangle_winding_is_consistent():
    """Every triangle's normal (v1-v0)x(v2-v0) at edge midpoints points the
    same way along its own edges (from the inside corner, f<T, to the
    outside one) for all 256 cases."""
    signs = set()
    for code, c in enumerate(_doc()["cases"]):
        inside = lambda p: (code >> (int(p[0]) + 2 * int(p[1]) + 4 * int(p[2]))) & 1
        for t in range(c["count"] // 3):
            tri = c["slots"][3 * t:3 * t + 3]
            mids = [sum(_edge_endpoints(s)) / 2.0 for s in tri]
            n = np.cross(mids[1] - mids[0], mids[2] - mids[0])
            score = 0.0
            for s in tri:
                a, b = _edge_endpoints(s)
                if not inside(a):
                    a, b = b, a
                score += float(np.dot(n, b - a))
            assert abs(score) > 1e-9, (code, t)
            signs.add(score > 0)
    assert len(signs) == 1


def test_device_table_matches_fixture():
    text = open(os.path.join(ROOT, "hpmc_amd", "csrc", "mc_tables.h")).read()
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{16})ull", text)]
    assert len(words) == 256
    for code, c in enumerate(_doc()["cases"]):
        w = words[code]
        assert w & 0xF == c["count"]
        for i, (sx, sy, sz, axis) in enumerate(c["slots"]):
            e = (w >> (4 + 4 * i)) & 0xF
            assert e >> 2 == axis
            sh = [sx, sy, sz]
            others = [q for q in range(3) if q != axis]
            assert (e & 1, (e >> 1) & 1) == (sh[others[0]], sh[others[1]])
        assert w >> (4 + 4 * c["count"]) == 0
