"""DRAT lines of the device ingestion (pfs_parse_dimacs with "proof"): on
every DIMACS fixture (messy texts with comments, duplicates, tautologies,
units, shrinking clauses; the self-refuting ones) and on the generated
texts, the lines are a valid RUP proof from the text's clauses
(tests/drat.py), a refuted text ends with the empty clause, and replaying
the lines on the text's clauses leaves exactly the resident formula: its
clauses and its root units.  The same proof followed by the simplify call's
own lines stays valid through the simplify (parse -> simplify)."""
import json
import os
import tempfile
from collections import Counter

import numpy as np
import pytest

from parafrost_amd import gen
from tests import drat
from tests.common import GOLDEN

pytestmark = pytest.mark.gpu
with open(os.path.join(GOLDEN, "dimacs_manifest.json")) as f:
    DMANIFEST = [e for e in json.load(f) if e["outcome"] != "error"]


def text_clauses(text: bytes):
    """The clauses of a DIMACS text as literal codes 2v + sign (comments,
    the header and a '%' ending skipped)."""
    out, cur = [], []
    for line in text.replace(b"\r", b"").split(b"\n"):
        t = line.strip()
        if not t or t.startswith(b"c") or t.startswith(b"p"):
            continue
        if t.startswith(b"%"):
            break
        for tok in t.split():
            if tok.startswith(b"c"):
                break
            x = int(tok)
            if x == 0:
                out.append(cur)
                cur = []
            else:
                cur.append(2 * abs(x) + (1 if x < 0 else 0))
    return out


def replay(clauses, proof):
    """The clause multiset left after the proof's additions and deletions."""
    live = Counter(tuple(sorted(set(c))) for c in clauses)
    for kind, lits in drat.parse(proof):
        key = tuple(sorted(set(lits)))
        if kind == "a":
            live[key] += 1
        else:
            live[key] -= 1
            assert live[key] >= 0
    return +live


def check_ingest(text, maxvar_hint=None):
    from parafrost_amd.simplifier import Simplifier
    cls = text_clauses(text)
    s = Simplifier(0, proof=1)
    try:
        info = s.parse_dimacs(text)
        proof = s.proof()
        trail = s.trail()
        res = s.result(2) if info["status"] != 0 else None
    finally:
        s.close()
    nv = max([maxvar_hint or 0, info["maxvar"]] + [l >> 1 for c in cls for l in c])
    ok, msg = drat.check(nv, cls, [], proof, expect_empty=info["status"] == 0)
    assert ok, msg
    if info["status"] == 0:
        return info
    left = replay(cls, proof)
    off = np.concatenate([[0], np.cumsum(res.sizes.astype(np.int64))])
    want = Counter(tuple(sorted(res.lits[off[i]:off[i + 1]].tolist())) for i in range(res.sizes.shape[0]))
    units = Counter((int(u),) for u in trail.tolist())
    got_cls = Counter({k: v for k, v in left.items() if len(k) > 1})
    got_units = {k for k in left if len(k) == 1}
    assert got_cls == want
    assert set(units) <= got_units and all(len(k) == 1 and (k[0] >> 1) in {u >> 1 for u in trail.tolist()}
                                           for k in got_units)
    return info


@pytest.mark.parametrize("entry", DMANIFEST, ids=[e["name"] for e in DMANIFEST])
def test_dimacs_fixture_proof(entry):
    with open(os.path.join(GOLDEN, entry["dimacs"]), "rb") as f:
        text = f.read()
    print("ingest proof", entry["name"], flush=True)
    info = check_ingest(text)
    assert (info["status"] == 0) == (entry["outcome"] == "unsat")


@pytest.mark.parametrize("name,make", [
    ("uniform3", lambda: gen.uniform_ksat(3000, 12780, 3, seed=5)),
    ("tseitin_units", lambda: gen.tseitin(6000, 150, 15, seed=6)),
])
def test_generated_text_proof_through_simplify(name, make):
    """Parse proof, then the simplify call's proof: together a valid proof
    from the text, ending at the simplified formula."""
    from parafrost_amd.simplifier import Simplifier
    cnf = make()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        with open(p, "rb") as f:
            text = f.read()
    cls = text_clauses(text)
    s = Simplifier(0, proof=1, phases=1, bce=1)
    try:
        info = s.parse_dimacs(text)
        p1 = s.proof()
        st = s.simplify()
        p2 = s.proof()
    finally:
        s.close()
    ok, msg = drat.check(info["maxvar"], cls, [], p1 + p2, expect_empty=st == 0)
    assert ok, msg
