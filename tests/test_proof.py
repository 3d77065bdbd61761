"""DRAT proof output of the oracle: against the reference's own proofs and
against the RUP checker of tests/drat.py (which must also reject planted
mistakes)."""
import random

import numpy as np
import pytest

from oracle import oracle
from tests import drat
from tests.common import clause_list, load_manifest, oracle_kw, read_golden, read_golden_proof

MANIFEST = load_manifest()


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
def test_oracle_proof_matches_reference(entry):
    """REF mode reproduces the reference simplifier's proof byte for byte."""
    inp, _ = read_golden(entry)
    r = oracle.simplify_state(inp, det=0, proof_en=1, **oracle_kw(entry["options"]))
    assert r.proof == read_golden_proof(entry)


@pytest.mark.parametrize("entry", MANIFEST[::3], ids=[e["name"] for e in MANIFEST[::3]])
def test_reference_proof_is_rup(entry):
    inp, out = read_golden(entry)
    ok, msg = drat.check(inp.maxvar, clause_list(inp), inp.units, read_golden_proof(entry),
                         expect_empty=out.status == 0)
    assert ok, msg


def _nontrivial_proof():
    for e in MANIFEST:
        if e["name"].startswith("ts_g2000.p1bce"):
            inp, _ = read_golden(e)
            return inp, read_golden_proof(e)
    raise AssertionError("fixture missing")


def _encode(lines):
    out = bytearray()
    for kind, lits in lines:
        out += kind.encode()
        for r in lits:
            while r > 127:
                out.append((r & 127) | 128)
                r >>= 7
            out.append(r)
        out.append(0)
    return bytes(out)


def test_checker_rejects_planted_mistakes():
    inp, proof = _nontrivial_proof()
    lines = drat.parse(proof)
    assert _encode(lines) == proof
    cls = clause_list(inp)
    ok, msg = drat.check(inp.maxvar, cls, inp.units, proof)
    assert ok, msg
    rng = random.Random(5)
    adds = [i for i, (k, l) in enumerate(lines) if k == "a" and len(l) >= 2]
    rejected = 0
    for i in rng.sample(adds, 8):
        bad = list(lines)
        kind, lits = bad[i]
        j = rng.randrange(len(lits))
        bad[i] = (kind, lits[:j] + lits[j + 1:])  # a dropped literal: no longer implied
        ok, _ = drat.check(inp.maxvar, cls, inp.units, _encode(bad))
        rejected += not ok
    assert rejected >= 6, rejected
    # a deletion of a clause that is not there
    ok, _ = drat.check(inp.maxvar, cls, inp.units, _encode(lines + [("d", (2, 4, 6))]))
    assert not ok
    # an empty clause on a satisfiable formula
    ok, _ = drat.check(inp.maxvar, cls, inp.units, _encode([("a", ())]))
    assert not ok


def test_checker_closed_form():
    """x1 & (x1 -> x2) & (x2 -> x3): (x3) and (x2 | x5) are RUP, (x4) is not."""
    cls = [[2], [3, 4], [5, 6]]
    assert drat.check(5, cls, [], _encode([("a", (6,)), ("a", (4, 10))]))[0]
    assert not drat.check(5, cls, [], _encode([("a", (8,))]))[0]
    # refutation: adding (-x3) makes the formula UNSAT, the empty clause is RUP
    assert drat.check(5, cls + [[7]], [], _encode([("a", ())]), expect_empty=True)[0]
