"""The DIMACS parser restatement (oracle/dimacs.py) against the reference.

* the reference's own parse + root BCP + shrinkTop (the simplifier input
  <case>.in.pfd.gz written by oracle/ref_harness.cpp) on the generated
  golden texts and on the hand-made parser fixtures (comments, blanks,
  CRLF, '+' signs, leading zeros, duplicates, tautologies, units, '%'),
  including the texts it refutes and the ones it rejects;
* known answers on tiny texts;
* planted mistakes in the restatement are caught by the fixtures.
"""
import json
import os
import sys
import tempfile
import types

import numpy as np
import pytest

from oracle import dimacs
from parafrost_amd import gen, pfd
from tests.common import GOLDEN

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

with open(os.path.join(GOLDEN, "dimacs_manifest.json")) as f:
    DMANIFEST = json.load(f)
GENERATED = [name for name, _, learnt in make_golden.cases() if not learnt]  # plain parses


def generated_text(name):
    cnf = {n: c for n, c, _ in make_golden.cases()}[name]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        with open(p, "rb") as f:
            return f.read()


def clause_sets(st):
    off = np.zeros(st.sizes.shape[0] + 1, np.int64)
    np.cumsum(st.sizes, out=off[1:])
    return [sorted(st.lits[off[k]:off[k + 1]].tolist()) for k in range(st.sizes.shape[0])]


def outcome(parse, text):
    try:
        st, _, _, cls, trail = parse(text)
    except dimacs.ParseError:
        return "error", None, None
    return ("unsat" if st == 0 else "ok"), cls, trail


def agrees(parse, text, entry_input, want):
    got, cls, trail = outcome(parse, text)
    if got != want:
        return False
    if got != "ok":
        return True
    inp = pfd.read(os.path.join(GOLDEN, entry_input))
    return [sorted(c) for c in cls] == clause_sets(inp) and sorted(trail) == sorted(inp.units.tolist())


@pytest.mark.parametrize("name", GENERATED)
def test_generated_texts_match_reference(name):
    assert agrees(dimacs.parse, generated_text(name), name + ".in.pfd.gz", "ok")


@pytest.mark.parametrize("entry", DMANIFEST, ids=[e["name"] for e in DMANIFEST])
def test_parser_fixtures_match_reference(entry):
    with open(os.path.join(GOLDEN, entry["dimacs"]), "rb") as f:
        text = f.read()
    assert agrees(dimacs.parse, text, entry.get("input"), entry["outcome"])


def test_fixture_outcomes_cover_every_branch():
    kinds = {e["outcome"] for e in DMANIFEST}
    assert kinds == {"ok", "unsat", "error"}


KAT = [
    (b"p cnf 3 2\n1 -2 0\n2 3 0\n", 2, [[2, 5], [4, 6]], []),
    # a parse-time unit shrinks a later clause to a unit
    (b"p cnf 3 2\n-1 0\n1 2 3 0\n", 2, [[4, 6]], [3]),
    (b"p cnf 3 2\n-1 0\n1 2 0\n", 2, [], [3, 4]),
    # duplicates dropped, tautology removed
    (b"p cnf 2 3\n1 1 2 0\n1 -1 0\n2 -1 0\n", 2, [[2, 4], [4, 3]], []),
    # root propagation through a chain satisfies everything
    (b"p cnf 3 3\n-1 2 0\n-2 3 0\n1 0\n", 2, [], [2, 4, 6]),
    # SATLIB ending, '+' sign, leading zeros, tabs, CRLF, comment after a 0
    (b"c x\r\np cnf 4 2\r\n+1\t02 0 c tail\r\n-3 4 0\r\n%\r\n0\r\n", 2, [[2, 4], [7, 8]], []),
    # the clause part ends at a NUL byte
    (b"p cnf 3 2\n1 2 0\n\x003 0\n", 2, [[2, 4]], []),
    # a clause spread over lines, a comment line between clauses
    (b"p cnf 3 2\n1\n2\n0\nc 1 2 0\n-3 1 0\n", 2, [[2, 4], [7, 2]], []),
    (b"p cnf 2 2\n1 0\n-1 0\n", 0, [], [2]),
    (b"p cnf 2 2\n0\n1 2 0\n", 0, [], []),
]


@pytest.mark.parametrize("k", range(len(KAT)))
def test_known_answers(k):
    text, st, cls, trail = KAT[k]
    r = dimacs.parse(text)
    assert r[0] == st
    if st:
        assert r[3] == cls
        assert sorted(r[4]) == sorted(trail)


@pytest.mark.parametrize("text", [b"p cnf 3 1\n1 4 0\n", b"p cnf 3 1\n1 x 0\n", b"1 2 0\n", b"p cnf 3 1\n1 2",
                                  b"p cnf 3 1\n1 2 0\n2 3 0\n", b"p cnf 0 1\n", b"p cnf 3 0\n", b"p cnf -3 1\n",
                                  b"p cnx 3 1\n", b"c only a comment\n"])
def test_rejected_texts(text):
    with pytest.raises(dimacs.ParseError):
        dimacs.parse(text)


MUTANTS = {
    "tautology kept": ("            elif m != s:\n                satisfied = True  # tautology\n", ""),
    "duplicate kept": ("            m = marks.get(v)\n", "            m = None\n"),
    "false literal kept": ("        out.append([x for x in cl if val(x) is None])\n", "        out.append(list(cl))\n"),
    "unit not assigned": ("            value[cl[0] >> 1] = cl[0]  # enqueueUnit\n", ""),
    "satisfied clause kept": ("        if any(val(x) for x in cl):\n            continue\n        out.append",
                              "        out.append"),
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(mutant):
    """Each slip, planted in a copy of the restatement, disagrees with the
    reference on at least one fixture."""
    src = open(dimacs.__file__).read()
    old, new = MUTANTS[mutant]
    assert old in src
    mod = types.ModuleType("dimacs_mutant")
    exec(compile(src.replace(old, new, 1), "dimacs_mutant", "exec"), mod.__dict__)
    cases = [(generated_text(n), n + ".in.pfd.gz", "ok") for n in GENERATED[:6]]
    for e in DMANIFEST:
        with open(os.path.join(GOLDEN, e["dimacs"]), "rb") as f:
            cases.append((f.read(), e.get("input"), e["outcome"]))
    assert any(not agrees(mod.parse, t, i, o) for t, i, o in cases)
