"""LUT fixture: structural invariants and the packed header agree with it."""

import json
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
LUTS = json.loads((ROOT / "isoext_amd/luts/mc_luts.json").read_text())
EDGES = [tuple(e) for e in LUTS["edge_corners"]]


def crossing(case):
    return sum(1 << e for e, (a, b) in enumerate(EDGES) if ((case >> a) & 1) != ((case >> b) & 1))


def tris(row):
    out = []
    for k in range(0, len(row), 3):
        if row[k] == -1:
            break
        out.append(tuple(row[k:k + 3]))
    return out


def test_edge_convention():
    # corner k at (k>>2, (k>>1)&1, k&1); every edge joins corners differing in one axis
    for a, b in EDGES:
        assert a < b and bin(a ^ b).count("1") == 1
    assert len(set(EDGES)) == 12


def test_table_shapes():
    assert len(LUTS["nagae"]) == 256 and all(len(r) == 15 for r in LUTS["nagae"])
    assert len(LUTS["lorensen"]) == 256 and all(len(r) == 12 for r in LUTS["lorensen"])


def test_each_case_uses_exactly_its_crossing_edges():
    for name, mx in (("nagae", 5), ("lorensen", 4)):
        for c, row in enumerate(LUTS[name]):
            ts = tris(row)
            assert len(ts) <= mx
            used = {e for t in ts for e in t}
            assert sum(1 << e for e in used) == crossing(c), (name, c)
            # no entries after the terminator
            n = 3 * len(ts)
            assert all(x == -1 for x in row[n:]), (name, c)


def test_complement_and_empty_cases():
    for name in ("nagae", "lorensen"):
        t = LUTS[name]
        assert tris(t[0]) == [] and tris(t[255]) == []
        for c in range(256):
            # complementary cases cross the same edges and emit the same count
            assert crossing(c) == crossing(255 - c)


def test_tables_differ_in_128_cases():
    n = sum(1 for c in range(256) if tris(LUTS["nagae"][c]) != tris(LUTS["lorensen"][c]))
    assert n == 128


def test_packed_header_matches_json():
    text = (ROOT / "isoext_amd/csrc/mc_luts.h").read_text()
    for name, key in (("kNagaeTris", "nagae"), ("kLorensenTris", "lorensen")):
        body = re.search(name + r"\[256\] = \{(.*?)\};", text, re.S).group(1)
        words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]+)ull", body)]
        assert len(words) == 256
        for c, w in enumerate(words):
            ts = tris(LUTS[key][c])
            assert w & 15 == len(ts)
            for k, t in enumerate(ts):
                for j in range(3):
                    assert (w >> (4 + 4 * (3 * k + j))) & 15 == t[j]
