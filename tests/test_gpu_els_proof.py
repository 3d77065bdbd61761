"""DRAT lines of the device equivalent-literal substitution
(pfs_decompose with "proof"): on every ELS fixture input the lines are a
valid RUP proof from the input (tests/drat.py), a refutation ends with the
empty clause, and replaying them on the input leaves exactly the resident
formula after the call.  A proof spanning parse -> ELS -> simplify, each
call's lines appended, is valid from the DIMACS text."""
import json
import os
import tempfile
from collections import Counter

import numpy as np
import pytest

from parafrost_amd import gen, pfd
from tests import drat
from tests.common import GOLDEN, clause_list
from tests.test_gpu_ingest_proof import replay, text_clauses

pytestmark = pytest.mark.gpu
with open(os.path.join(GOLDEN, "els_manifest.json")) as f:
    EMANIFEST = json.load(f)


def resident_clauses(res):
    off = np.concatenate([[0], np.cumsum(res.sizes.astype(np.int64))])
    return Counter(tuple(sorted(res.lits[off[i]:off[i + 1]].tolist())) for i in range(res.sizes.shape[0]))


@pytest.mark.parametrize("entry", EMANIFEST, ids=[e["name"] for e in EMANIFEST])
def test_els_fixture_proof(entry):
    from parafrost_amd.simplifier import Simplifier
    inp = pfd.read(os.path.join(GOLDEN, entry["input"]))
    print("els proof", entry["name"], flush=True)
    s = Simplifier(0, proof=1)
    try:
        s.upload_state(inp)
        info = s.decompose(entry["rounds"] or 0)
        proof = s.proof()
        res = s.result(2) if info["status"] != 0 else None
        trail = s.trail() if info["status"] != 0 else None
    finally:
        s.close()
    cls = clause_list(inp)
    ok, msg = drat.check(inp.maxvar, cls, inp.units.tolist(), proof, expect_empty=info["status"] == 0)
    assert ok, msg
    if info["status"] == 0:
        return
    left = replay(cls + [[int(u)] for u in inp.units.tolist()], proof)
    got = Counter({k: v for k, v in left.items() if len(k) > 1})
    assert got == resident_clauses(res)
    assert {k[0] for k in left if len(k) == 1} <= set(trail.tolist())


@pytest.mark.parametrize("seed,anti", [(21, 0), (22, 8), (23, 40)])
def test_parse_els_simplify_proof(seed, anti):
    from parafrost_amd.simplifier import Simplifier
    cnf = gen.equivalences(3000, 9000, 150, 5, seed=seed, anti=anti, units=4)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "e.cnf")
        gen.write_dimacs(cnf, p)
        with open(p, "rb") as f:
            text = f.read()
    cls = text_clauses(text)
    s = Simplifier(0, proof=1, phases=1, bce=1)
    try:
        i1 = s.parse_dimacs(text)
        p1 = s.proof()
        st = 0
        p2 = p3 = b""
        if i1["status"] != 0:
            i2 = s.decompose(3)
            p2 = s.proof()
            if i2["status"] != 0:
                st = s.simplify()
                p3 = s.proof()
    finally:
        s.close()
    ok, msg = drat.check(i1["maxvar"], cls, [], p1 + p2 + p3, expect_empty=st == 0)
    assert ok, msg
