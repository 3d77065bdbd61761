"""Parity of the device DIMACS ingestion (pfs_parse_dimacs) with the
reference's parser + root BCP + shrinkTop.  Integer/byte work: bit-exact
on the clause order, the clause literal sets (the resident formula keeps
literals sorted, as the simplifier's extraction does) and the trail's set.

* every golden text (generated and hand-made) parses to the reference's
  simplifier input, and simplifying it gives the reference's output;
* texts the reference refutes report UNSAT; texts it rejects are rejected;
* edge cases against the oracle restatement;
* full-size Tseitin text (2,000,000 gates): the resident formula equals
  the host-prepared one.
"""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle import dimacs
from parafrost_amd import gen, pfd
from parafrost_amd.simplifier import PfsError, Simplifier
from tests.common import GOLDEN, diff, load_manifest

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu
with open(os.path.join(GOLDEN, "dimacs_manifest.json")) as f:
    DMANIFEST = json.load(f)
# the fixtures whose input is the plain parse of the generated text (the
# learnt-clause fixtures hand some parsed clauses over as learnt)
MANIFEST = [e for e in load_manifest() if not e.get("learnt")]


def generated_text(name):
    cnf = {n: c for n, c, _ in make_golden.cases()}[name]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        with open(p, "rb") as f:
            return f.read()


def sets_of(sizes, lits):
    off = np.zeros(sizes.shape[0] + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    return [sorted(lits[off[k]:off[k + 1]].tolist()) for k in range(sizes.shape[0])]


def resident(s):
    r = s.result(2)
    return r, sets_of(r.sizes, r.lits), sorted(s.trail().tolist())


def check_input(s, info, inp):
    r, cls, trail = resident(s)
    assert info["maxvar"] == inp.maxvar
    assert cls == sets_of(inp.sizes, inp.lits)
    assert trail == sorted(inp.units.tolist())
    assert np.array_equal(r.vstate, inp.vstate)
    assert info["clauses"] == inp.sizes.shape[0] and info["trail"] == inp.units.shape[0]


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
def test_generated_text_parses_and_simplifies_like_reference(entry):
    name = entry["name"].split(".")[0]
    print("dimacs golden", entry["name"], flush=True)
    inp, out = pfd.read(os.path.join(GOLDEN, entry["input"])), pfd.read(os.path.join(GOLDEN, entry["output"]))
    s = Simplifier(0, **entry["options"])
    try:
        info = s.parse_dimacs(generated_text(name))
        check_input(s, info, inp)
        g = s.result(s.simplify())
    finally:
        s.close()
    d = diff(g, out)
    assert not d, d


@pytest.mark.parametrize("entry", DMANIFEST, ids=[e["name"] for e in DMANIFEST])
def test_parser_fixture_like_reference(entry):
    print("dimacs fixture", entry["name"], flush=True)
    with open(os.path.join(GOLDEN, entry["dimacs"]), "rb") as f:
        text = f.read()
    s = Simplifier(0, **entry["options"])
    try:
        if entry["outcome"] == "error":
            with pytest.raises(PfsError, match="BAD_INPUT"):
                s.parse_dimacs(text)
            return
        info = s.parse_dimacs(text)
        if entry["outcome"] == "unsat":
            assert info["status"] == 0
            assert s.simplify() == 0
            return
        check_input(s, info, pfd.read(os.path.join(GOLDEN, entry["input"])))
        g = s.result(s.simplify())
    finally:
        s.close()
    d = diff(g, pfd.read(os.path.join(GOLDEN, entry["output"])))
    assert not d, d


EDGE = [
    b"p cnf 3 2\n1 -2 0\n2 3 0",  # no final newline
    b"p cnf 3 2\n-1 0\n1 2 3 0\n",
    b"p cnf 2 3\n1 1 2 0\n1 -1 0\n2 -1 0\n",
    b"p cnf 3 3\n-1 2 0\n-2 3 0\n1 0\n",
    b"c x\r\np cnf 4 2\r\n+1\t02 0 c tail\r\n-3 4 0\r\n%\r\n0\r\n",
    b"p cnf 3 2\n1 2 0\n\x003 0\n",
    b"p cnf 3 2\n1\n2\n0\nc 1 2 0\n-3 1 0\n",
    b"p cnf 3 3\nc one % two\n1 2 0\nc c c 5 0\n-0\n",  # '%' and 'c' inside comments; -0 ends an empty clause
    b"p cnf 5 4\n" + b"c " + b"z" * 5000 + b"\n1 2 0\n" + b" " * 9000 + b"3 -4 5 0\n2 -2 0\n-5 0\n",
    b"p cnf 3 1 1 2 0\n",  # a clause on the header line
]


@pytest.mark.parametrize("k", range(len(EDGE)))
def test_edge_texts_match_oracle(k):
    text = EDGE[k]
    st, mv, _, cls, trail = dimacs.parse(text)
    s = Simplifier(0)
    try:
        info = s.parse_dimacs(text)
        assert info["status"] == st
        if st:
            r, got, tr = resident(s)
            assert got == [sorted(c) for c in cls]
            assert tr == sorted(trail)
    finally:
        s.close()


@pytest.mark.parametrize("text", [b"p cnf 3 1\n1-2 0\n", b"p cnf 3 2\n1 2 0c\n", b"p cnf 3 1\n1 99999999999 0\n",
                                  b"p cnf 3 1\np cnf 3 1\n1 0\n"])
def test_rejected_by_device_lexer(text):
    """Numbers and markers must be separated by blanks (the reference also
    splits '1-2' and '0c'); a second header and oversized numbers are
    rejected."""
    s = Simplifier(0)
    try:
        with pytest.raises(PfsError, match="BAD_INPUT"):
            s.parse_dimacs(text)
    finally:
        s.close()


def test_full_tseitin_text():
    """The 2,000,000-gate Tseitin circuit as text: the device ingestion
    gives the host-prepared formula (units propagated, satisfied clauses
    removed) clause for clause."""
    raw = gen.tseitin(2_000_000, 5000, 1000, seed=0x5EED0002, sat_outputs=True)
    prep, refuted = gen.prepare(raw)
    assert not refuted
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.cnf")
        gen.write_dimacs(raw, p)
        with open(p, "rb") as f:
            text = f.read()
    print("tseitin text bytes", len(text), flush=True)
    s = Simplifier(0, keep_input=True)
    try:
        info = s.parse_dimacs(text)
        r = s.result(2)
        trail = s.trail()
    finally:
        s.close()
    assert info["read_clauses"] == raw.nclauses
    assert np.array_equal(r.sizes, np.diff(prep.offsets).astype(np.uint32))
    # sort the literals of each clause of the prepared formula
    cid = np.repeat(np.arange(prep.nclauses), np.diff(prep.offsets).astype(np.int64))
    order = np.lexsort((prep.lits, cid))
    assert np.array_equal(r.lits, prep.lits[order])
    assert np.array_equal(np.sort(trail), np.sort(prep.trail))


def test_refuted_circuit_text():
    """Asserted outputs with random values usually contradict the circuit:
    the device propagation refutes the text exactly when the host
    preparation does."""
    raw = gen.tseitin(200_000, 500, 400, seed=7)
    _, refuted = gen.prepare(raw)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.cnf")
        gen.write_dimacs(raw, p)
        with open(p, "rb") as f:
            text = f.read()
    s = Simplifier(0)
    try:
        info = s.parse_dimacs(text)
    finally:
        s.close()
    assert (info["status"] == 0) == refuted


def test_parse_then_rewind():
    text = generated_text("ts_g2000")
    s = Simplifier(0, keep_input=True, phases=1, bce=1)
    try:
        s.parse_dimacs(text)
        a = s.result(s.simplify())
        s.rewind()
        b = s.result(s.simplify())
    finally:
        s.close()
    assert not diff(a, b)
