"""Pins the oracle's cut constants and chi-squared tables to the reference.

tests/golden/reference_constants.json is extracted from the reference
sources by tests/golden/extract_reference_constants.py (text only).  Two
checks a plausible mistake in the restatement fails:

* every oracle function that cites reference lines (in the comment above it
  and in comments inside it) holds the same multiset of non-integer numeric
  literals as those lines, except the differences listed in EXPLAINED, each
  with its reason: a mistyped or dropped threshold fails;
* each chi-squared table function of the oracle, probed at every layer
  combination of the reference table's domain just below, at and just above
  the threshold the reference code returns there, gives the reference's
  answer: a wrong layer number, a shadowed or missing branch, or a swapped
  '<' / '<=' fails.
"""
import collections
import ctypes as C
import itertools
import json
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "reference_constants.json")
ORACLE = os.path.join(ROOT, "oracle")
NUM = re.compile(r"(?<![\w.])((?:\d+\.\d*|\.\d+)(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)[fF]?(?![\w.])")

# (oracle file, first cited range of the function) -> (literals only the reference lines hold, literals only
# the oracle function holds, why).  Values are float32 literals with their counts.
EXPLAINED = {
    ("orc_ls_t3.cpp", "SDL/Segment.h:177-200"): (
        {0.26: 4, 0.16: 2, 0.18: 33, 0.4: 48}, {},
        "moduleGapSize_seg's nested if/else tree (SDL/Segment.h:202-262) is a lookup table in the oracle; the "
        "table's values are the same set, each written once"),
    ("orc_md.cpp", "SDL/MiniDoublet.h:286-337"): (
        {0.26: 1, 0.16: 2, 0.18: 35, 0.4: 50}, {},
        "moduleGapSize's nested if/else tree is a lookup table in the oracle (same set of values)"),
    ("orc_t5.cpp", "SDL/Quintuplet.h:218-298"): (
        {0.04725: 1, 0.09461: 1}, {},
        "the (1,2,3,4,x) branch is shadowed by the (1,2,3) branch before it and never taken; the table probe "
        "below pins that"),
    ("orc_t5.cpp", "SDL/Quintuplet.h:1223-1320"): (
        {0.5: 5, 0.0: 1}, {},
        "the five 0.5f * float(M_PI) angle cases are one helper shared with computeChiSquared "
        "(angleM_of); the 0. is the reference's unused default of a local"),
    ("orc_t5.cpp", "SDL/Quintuplet.h:1322-1367"): (
        {0.5: 5}, {}, "the same angle helper"),
    ("orc_t5.cpp", "SDL/Quintuplet.h:1744-2028"): (
        {1.0: 4, 0.0: 1, 2.0: 2, 3.0: 1}, {},
        "the z/rt window, multiple-scattering and beta-correction expressions shared with the barrel "
        "selector are helpers called from both"),
    ("orc_t5.cpp", "SDL/Quintuplet.h:2030-2298"): (
        {1.0: 2, 0.0: 1}, {}, "the same shared helpers"),
    ("orc_tc.cpp", "SDL/PixelTriplet.h:615-734"): (
        {2.0: 3, 0.0: 3, 1.0: 6}, {},
        "the four radius-criterion variants share one inverse-radius interval helper"),
    ("orc_tc.cpp", "SDL/PixelTriplet.h:826-1024"): (
        {5.0: 2, 0.009: 2, 0.01: 1, 0.15: 1}, {},
        "the three 'runChiSquaredCuts and pixelSegmentPt < 5.0f' tests are one flag; computePT3RPhiChiSquared's "
        "sigmas (SDL/PixelTriplet.h:419-487) are the shared sigmasForRegression helper, whose constants are "
        "checked with SDL/Quintuplet.h:1136-1221"),
    ("orc_t5.cpp", "SDL/Event.cc:903-991"): (
        {}, {0.0: 1}, "a result buffer cleared with 0.f (the reference clears its device buffers with memset)"),
    ("orc_tc.cpp", "SDL/Event.cc:1009-1146"): (
        {}, {0.0: 1}, "a result buffer cleared with 0.f (the reference clears its device buffers with memset)"),
    ("orc_tc.cpp", "SDL/Event.cc:784-901"): (
        {}, {0.0: 1}, "a result buffer cleared with 0.f (the reference clears its device buffers with memset)"),
}
IGNORED = {
    ("orc_geometry.cpp", "SDL/ModuleConnectionMap.cc:9-50"): "geometry file reader: no cut constants",
    ("orc_geometry.cpp", "SDL/TiltedGeometry.cc:5-49"): "geometry file reader: no cut constants",
}


def f32(x):
    return struct.unpack("<f", struct.pack("<f", float(x)))[0]


def literals(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return collections.Counter(f32(m.group(1)) for m in NUM.finditer(text))


def line_ranges(comment, f=None):
    """Cited ranges; a ':a-b' without a file means the file cited last before it (`f` carries that in)."""
    out = []
    for file_, a, b in re.findall(r"(SDL/[\w.]+)?:(\d+)-(\d+)", comment):
        f = file_ or f
        if f:
            out.append(f"{f}:{a}-{b}")
    return out


def oracle_functions():
    """(file, cited ranges, body) of every oracle function with a citation comment above it."""
    out = []
    for name in sorted(os.listdir(ORACLE)):
        if not name.endswith(".cpp"):
            continue
        src = open(os.path.join(ORACLE, name)).read().split("\n")
        for i, line in enumerate(src):
            if not re.match(r"\s*// SDL/\S+:\d+-\d+", line):
                continue
            j = i + 1
            while j < len(src) and "{" not in src[j]:
                j += 1
            depth, body = 0, []
            for k in range(j, len(src)):
                body.append(src[k])
                depth += src[k].count("{") - src[k].count("}")
                if depth <= 0 and "}" in src[k]:
                    break
            cites = line_ranges(line)
            for b in body:  # citations inside the body extend the function's reference lines
                if "//" in b:
                    last = cites[-1].rsplit(":", 1)[0]
                    cites += [c for c in line_ranges(b[b.index("//"):], last) if c not in cites]
            out.append((name, cites, "\n".join(body)))
    return out


@pytest.fixture(scope="module")
def fixture():
    return json.load(open(FIXTURE))


def test_every_cited_function_is_checked(fixture):
    funcs = oracle_functions()
    assert len(funcs) > 50
    for name, cites, _ in funcs:
        for c in cites:
            assert c in fixture["ranges"], f"{name}: {c} missing from the fixture (rerun the extraction)"


def reference_literals(fixture, cites):
    """The literals of the union of the cited line ranges (overlapping citations count each line once)."""
    lines = {}
    for c in cites:
        f, r = c.rsplit(":", 1)
        a, b = (int(x) for x in r.split("-"))
        lines.setdefault(f, set()).update(range(a, b + 1))
    out = collections.Counter()
    for f, ns in lines.items():
        per_line = fixture["literals"].get(f, {})
        for n in ns:
            for k, cnt in per_line.get(str(n), {}).items():
                out[f32(k)] += cnt
    return out


def test_oracle_literals_match_reference(fixture):
    """A mistyped threshold would change a function's literal multiset."""
    bad, used = [], set()
    for name, cites, body in oracle_functions():
        key = (name, cites[0])
        if key in IGNORED:
            continue
        ref = reference_literals(fixture, cites)
        orc = literals(body)
        only_ref, only_orc = ref - orc, orc - ref
        exp = EXPLAINED.get(key)
        if exp is not None:
            used.add(key)
            want_ref = collections.Counter({f32(k): n for k, n in exp[0].items()})
            want_orc = collections.Counter({f32(k): n for k, n in exp[1].items()})
            if only_ref != want_ref or only_orc != want_orc:
                bad.append(f"{name} {cites[0]}: reference-only {dict(only_ref)}, oracle-only {dict(only_orc)}, "
                           f"explained {dict(want_ref)} / {dict(want_orc)}")
        elif only_ref or only_orc:
            bad.append(f"{name} {cites[0]}: reference-only {dict(only_ref)}, oracle-only {dict(only_orc)}")
    assert not bad, "\n".join(bad)
    assert used == set(EXPLAINED), f"stale explanations: {set(EXPLAINED) - used}"


def _tables():
    from oracle import orc
    L = orc.lib()
    L.orc_pixel_cut_table.argtypes = [C.c_char_p, C.c_void_p, C.c_float]
    L.orc_pixel_cut_table.restype = C.c_int
    L.orc_t5_cut_table.argtypes = [C.c_void_p, C.c_float]
    L.orc_t5_cut_table.restype = C.c_int

    def call(name, layers, value):
        arr = np.ascontiguousarray(layers, np.int32)
        p = arr.ctypes.data_as(C.c_void_p)
        if name == "passChiSquaredConstraint":
            return L.orc_t5_cut_table(p, value)
        r = L.orc_pixel_cut_table(name.encode(), p, value)
        assert r >= 0, name
        return r
    return call


@pytest.mark.parametrize("name", ["passChiSquaredConstraint", "passPT3RZChiSquaredCuts", "passPT3RPhiChiSquaredCuts",
                                  "passPT3RPhiChiSquaredInwardsCuts", "passPT5RZChiSquaredCuts",
                                  "passPT5RPhiChiSquaredCuts", "passPT5RPhiChiSquaredInwardsCuts"])
def test_chi2_table_matches_reference(fixture, name):
    call = _tables()
    t = fixture["tables"][name]
    cases = {tuple(c[:t["arity"]]): (c[t["arity"]], c[t["arity"] + 1]) for c in t["cases"]}
    n_cmp = 0
    for tup in itertools.product(*t["domain"]):
        op, val = cases.get(tup, ("const", t["default"]))
        if op == "const":
            for x in (0.0, 1.0, 1e9):
                assert bool(call(name, tup, x)) == val, (name, tup, x)
            continue
        n_cmp += 1
        thr = np.float32(val)
        below, above = np.nextafter(thr, np.float32(-np.inf)), np.nextafter(thr, np.float32(np.inf))
        assert call(name, tup, float(below)) == 1, (name, tup, "below", thr)
        assert call(name, tup, float(thr)) == (1 if op == "<=" else 0), (name, tup, "at", thr, op)
        assert call(name, tup, float(above)) == 0, (name, tup, "above", thr)
    assert n_cmp == len(cases) and n_cmp >= 11
