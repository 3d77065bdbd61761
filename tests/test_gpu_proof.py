"""DRAT proof output of the HIP simplifier.

* golden fixtures: the GPU proof equals the reference simplifier's proof
  (runs of unit lines compared as sets: parallel propagation claims the
  same units in another order) and passes the RUP checker;
* seeded formulas of every family under several option sets: the GPU proof
  equals the oracle's DET-mode proof the same way, and is RUP; refutations
  end with an RUP empty clause.
"""
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests import drat
from tests.common import (clause_list, load_manifest, oracle_kw, proof_lines_canonical, read_golden,
                          read_golden_proof)
from tests.fuzz_cases import case

pytestmark = pytest.mark.gpu
MANIFEST = load_manifest()


def _run(maxvar, offsets, lits, trail=None, flags=None, lbd=None, **opts):
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0, proof=1, **opts)
    try:
        s.upload(maxvar, offsets, lits, flags, lbd, None, None, None, trail)
        return s.result(s.simplify())
    finally:
        s.close()


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
def test_golden_proof(entry):
    from parafrost_amd.simplifier import simplify_state
    inp, out = read_golden(entry)
    print("golden proof", entry["name"], flush=True)
    g = simplify_state(inp, proof=1, **entry["options"])
    assert g.status == out.status
    ref = read_golden_proof(entry)
    if out.status != 0:
        assert proof_lines_canonical(g.proof) == proof_lines_canonical(ref)
    ok, msg = drat.check(inp.maxvar, clause_list(inp), inp.units, g.proof, expect_empty=out.status == 0)
    assert ok, msg


@pytest.mark.parametrize("i", range(80))
def test_fuzz_proof(i):
    cnf, flags, lbd, ro, oo = case(i)
    print("fuzz proof", i, ro, cnf.nclauses, flush=True)
    g = _run(cnf.maxvar, cnf.offsets, cnf.lits, None, flags, lbd, **ro)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, flags, lbd, det=1, proof_en=1, **oo)
    assert g.status == r.status
    if r.status != 0:
        assert proof_lines_canonical(g.proof) == proof_lines_canonical(r.proof)
    off = cnf.offsets.tolist()
    cls = [cnf.lits[off[k]:off[k + 1]].tolist() for k in range(cnf.nclauses)]
    ok, msg = drat.check(cnf.maxvar, cls, [], g.proof, expect_empty=r.status == 0)
    assert ok, msg


@pytest.mark.parametrize("family", ["tseitin", "heavytail"])
def test_family_proof(family):
    cnf = (gen.prepare(gen.tseitin(6000, 150, 15, seed=311))[0] if family == "tseitin"
           else gen.heavytail(5000, 20000, seed=411))
    print("family proof", family, cnf.nclauses, flush=True)
    g = _run(cnf.maxvar, cnf.offsets, cnf.lits, cnf.trail, phases=2, bce=1)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1, proof_en=1,
                        vstate=_vstate(cnf), phases=2, bce_en=1)
    assert g.status == r.status
    assert proof_lines_canonical(g.proof) == proof_lines_canonical(r.proof)
    off = cnf.offsets.tolist()
    cls = [cnf.lits[off[k]:off[k + 1]].tolist() for k in range(cnf.nclauses)]
    ok, msg = drat.check(cnf.maxvar, cls, cnf.trail.tolist(), g.proof, expect_empty=r.status == 0)
    assert ok, msg


def _vstate(cnf):
    import numpy as np
    vs = np.zeros(cnf.maxvar + 1, np.uint8)
    vs[cnf.trail >> 1] = 2
    return vs
