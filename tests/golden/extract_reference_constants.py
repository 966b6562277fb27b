"""Write tests/golden/reference_constants.json from the reference sources.

Run in a container where the reference tree is present (it is read as text;
nothing of it is compiled, imported or run):

    python tests/golden/extract_reference_constants.py /root/reference

Two kinds of data are written:

* "ranges" and "literals": every reference line range a comment of the
  oracle cites ("// SDL/File.h:a-b ..."), and for each line of those ranges
  the multiset of its non-integer numeric literals (comments stripped), as
  float32 values.  A mistyped threshold in the oracle changes its function's
  multiset.
* "tables": the chi-squared cut tables of SDL/Quintuplet.h and
  SDL/PixelTriplet.h, evaluated from the reference code by a small
  interpreter of their if / else-if / return structure: for every tuple of
  layer numbers over the values the conditions name (and one value none of
  them names), either the comparison and threshold the reference returns
  ("<" or "<=", float32 value) or the constant it returns (stored: the
  domain, the constant most tuples return, and every other tuple).  This fixes the
  branch structure (which layer combinations reach which threshold, including
  branches an earlier condition shadows) and the comparison operators.
"""
from __future__ import annotations

import itertools
import json
import os
import re
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle")
OUT = os.path.join(HERE, "reference_constants.json")

NUM = re.compile(r"(?<![\w.])((?:\d+\.\d*|\.\d+)(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)[fF]?(?![\w.])")
CITE = re.compile(r"^\s*// (SDL/\S+?):(\d+)-(\d+)")

# chi-squared tables: (file, name, number of layer arguments)
TABLES = [("SDL/Quintuplet.h", "passChiSquaredConstraint", 5),
          ("SDL/PixelTriplet.h", "passPT3RZChiSquaredCuts", 3),
          ("SDL/PixelTriplet.h", "passPT3RPhiChiSquaredCuts", 3),
          ("SDL/PixelTriplet.h", "passPT3RPhiChiSquaredInwardsCuts", 3),
          ("SDL/PixelTriplet.h", "passPT5RZChiSquaredCuts", 5),
          ("SDL/PixelTriplet.h", "passPT5RPhiChiSquaredCuts", 5),
          ("SDL/PixelTriplet.h", "passPT5RPhiChiSquaredInwardsCuts", 5)]


def f32(x: str) -> float:
    return struct.unpack("<f", struct.pack("<f", float(x)))[0]


def strip_comments(s: str) -> str:
    """Comments removed, line breaks kept (so line numbers stay)."""
    s = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def literals(text: str) -> dict:
    out = {}
    for m in NUM.finditer(strip_comments(text)):
        k = repr(f32(m.group(1)))
        out[k] = out.get(k, 0) + 1
    return out


def line_ranges(comment: str, f=None) -> list:
    """The 'SDL/File:a-b' ranges a comment cites; a ':a-b' without a file means the file cited last before it
    (in this comment, or in the oracle source above it when `f` carries that)."""
    out = []
    for file_, a, b in re.findall(r"(SDL/[\w.]+)?:(\d+)-(\d+)", comment):
        f = file_ or f
        if f:
            out.append((f, int(a), int(b)))
    return out


def cited_ranges() -> list:
    """Every reference range a comment of the oracle cites."""
    cites = []
    for name in sorted(os.listdir(ORACLE)):
        if name.endswith(".cpp"):
            last = None
            for line in open(os.path.join(ORACLE, name)):
                if "//" in line:
                    got = line_ranges(line[line.index("//"):], last)
                    cites += got
                    if got:
                        last = got[-1][0]
    return sorted(set(cites))


# ---- a small interpreter of the chi-squared table functions
TOKEN = re.compile(r"\s*(==|<=|>=|&&|\|\||[(){};<>=!,+*-]|\w+(?:\.\w*)?|\S)")


def tokenize(s):
    return [t for t in TOKEN.findall(strip_comments(s)) if t.strip()]


def function_body(src: str, name: str) -> str:
    i = src.index(name + "(")
    i = src.index("{", src.index(")", i))
    depth, j = 0, i
    while True:
        depth += src[j] == "{"
        depth -= src[j] == "}"
        j += 1
        if depth == 0:
            return src[i:j]


class Parser:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, want=None):
        tok = self.t[self.i]
        if want is not None and tok != want:
            raise SyntaxError(f"expected {want!r}, got {tok!r} at token {self.i}")
        self.i += 1
        return tok

    def block(self):
        self.take("{")
        stmts = []
        while self.peek() != "}":
            s = self.stmt()
            if s is not None:
                stmts.append(s)
        self.take("}")
        return ("block", stmts)

    def stmt(self):
        tok = self.peek()
        if tok == "{":
            return self.block()
        if tok == "if":
            self.take("if")
            self.take("(")
            cond = self.cond()
            self.take(")")
            then = self.stmt()
            other = None
            if self.peek() == "else":
                self.take("else")
                other = self.stmt()
            return ("if", cond, then, other)
        if tok == "return":
            self.take("return")
            toks = []
            while self.peek() != ";":
                toks.append(self.take())
            self.take(";")
            if toks in (["true"], ["false"]):
                return ("const", toks[0] == "true")
            if len(toks) == 3 and toks[1] in ("<", "<="):
                return ("cmp", toks[1], f32(toks[2].rstrip("fF")))
            raise SyntaxError(f"unexpected return {' '.join(toks)}")
        # a declaration (the layer numbers): skipped
        while self.take() != ";":
            pass
        return None

    def cond(self):
        terms = [self.term()]
        while self.peek() in ("and", "&&"):
            self.take()
            terms.append(self.term())
        return terms

    def term(self):
        if self.peek() == "(":
            self.take("(")
            t = self.cond()
            self.take(")")
            return ("all", t)
        var = self.take()
        self.take("==")
        return (var, int(self.take()))


def holds(cond, env):
    for t in cond:
        if t[0] == "all":
            if not holds(t[1], env):
                return False
        elif env[t[0]] != t[1]:
            return False
    return True


def run(node, env):
    """The value of the first return reached, or None when control falls off the node."""
    kind = node[0]
    if kind == "block":
        for s in node[1]:
            r = run(s, env)
            if r is not None:
                return r
        return None
    if kind == "if":
        if holds(node[1], env):
            return run(node[2], env)
        return run(node[3], env) if node[3] is not None else None
    return node


def conditions(node, out):
    if node[0] == "block":
        for s in node[1]:
            conditions(s, out)
    elif node[0] == "if":
        stack = list(node[1])
        while stack:
            t = stack.pop()
            if t[0] == "all":
                stack += t[1]
            else:
                out.setdefault(t[0], set()).add(t[1])
        conditions(node[2], out)
        if node[3] is not None:
            conditions(node[3], out)


def table(ref_dir, file_, name, arity):
    src = open(os.path.join(ref_dir, file_)).read()
    body = function_body(src, name)
    tree = Parser(tokenize(body)).block()
    names = [f"layer{k}" for k in range(1, arity + 1)]
    seen = {}
    conditions(tree, seen)
    domain = [sorted(seen.get(n, set()) | {0}) for n in names]  # 0: a value no condition names
    results = {}
    for tup in itertools.product(*domain):
        r = run(tree, dict(zip(names, tup)))
        if r is None:
            raise SyntaxError(f"{name}: control falls off the function for {tup}")
        results[tup] = [r[1], r[2]] if r[0] == "cmp" else ["const", r[1]]
    consts = [v[1] for v in results.values() if v[0] == "const"]
    default = max(set(consts), key=consts.count)
    cases = [list(t) + v for t, v in results.items() if v != ["const", default]]
    line = src[:src.index(name + "(")].count("\n") + 1
    # every tuple of the domain's product not listed in cases returns `default`
    return {"ref": f"{file_}:{line}", "arity": arity, "domain": domain, "default": default, "cases": cases}


def main(argv):
    ref_dir = argv[1] if len(argv) > 1 else "/root/reference"
    data = {"ranges": [], "literals": {}, "tables": {}}
    for f, a, b in cited_ranges():
        data["ranges"].append(f"{f}:{a}-{b}")
        text = strip_comments(open(os.path.join(ref_dir, f)).read()).split("\n")  # line numbers kept
        per_line = data["literals"].setdefault(f, {})
        for n in range(a, b + 1):
            lits = literals(text[n - 1])
            if lits:
                per_line[str(n)] = lits
    for f, name, arity in TABLES:
        data["tables"][name] = table(ref_dir, f, name, arity)
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=0, sort_keys=True)
    print(f"wrote {OUT}: {len(data['ranges'])} cited ranges, "
          f"{sum(len(t['cases']) for t in data['tables'].values())} table thresholds")


if __name__ == "__main__":
    main(sys.argv)
