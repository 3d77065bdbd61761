"""Equisatisfiability of the GPU outputs, checked by the test CDCL solver
(tests/cdcl): the output (clauses plus every root unit) is solved; a model
is extended over the GPU's eliminated-variable stack (MODEL::extend,
src/gpu/model.cpp:101-162) and must satisfy the original formula; a refuted
output needs a refuted input.  Learnt clauses of an input are redundant by
contract and are left out of the original formula.  Random 3-SAT at the
satisfiability threshold is left out (beyond the test solver); the
heavy-tailed formulas are refuted (their head variables meet in binary
clauses of every sign), so their check is refutation on both sides."""
import numpy as np
import pytest

from parafrost_amd import gen
from tests import cdcl
from tests.common import load_manifest, read_golden
from tests.fuzz_cases import case as fuzz_case

pytestmark = pytest.mark.gpu
HARD = ("u3_n300.", "u3_n2000.", "lt_u3_n2000.")
# the learnt-clause fixtures plant random learnt clauses (not implied): left out
MANIFEST = [e for e in load_manifest() if not e["name"].startswith(HARD) and not e.get("learnt")]


def gpu(maxvar, offsets, lits, trail=None, flags=None, lbd=None, vstate=None, **opts):
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0, **opts)
    try:
        s.upload(maxvar, offsets, lits, flags, lbd, vstate, None, None, trail)
        return s.result(s.simplify())
    finally:
        s.close()


def check(maxvar, sizes_or_offsets, lits, units, flags, g, offsets=False):
    if offsets:
        sizes = np.diff(np.asarray(sizes_or_offsets, np.uint64))
    else:
        sizes = np.asarray(sizes_or_offsets)
    fl = np.zeros(sizes.shape[0], np.uint8) if flags is None else flags
    off, ol = cdcl.original_clauses(sizes, fl, lits)
    return cdcl.check_equisat(maxvar, off, ol, units, g.status, g.sizes, g.lits, g.units, g.resolved)


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
def test_golden_inputs(entry):
    from parafrost_amd.simplifier import simplify_state
    inp, _ = read_golden(entry)
    print("equisat golden", entry["name"], flush=True)
    g = simplify_state(inp, **entry["options"])
    assert check(inp.maxvar, inp.sizes, inp.lits, inp.units, inp.flags, g)


@pytest.mark.parametrize("i", range(60))
def test_fuzz_formulas(i):
    cnf, flags, lbd, ro, _ = fuzz_case(i)
    if flags is not None:
        pytest.skip("its learnt clauses are drawn at random, not implied: no equisatisfiability to check")
    prep, refuted = (cnf, False) if flags is not None else gen.prepare(cnf)
    if refuted:
        pytest.skip("refuted at the root")
    g = gpu(prep.maxvar, prep.offsets, prep.lits, prep.trail, flags, lbd, **ro)
    assert check(prep.maxvar, prep.offsets, prep.lits, prep.trail, flags, g, offsets=True)


FAMILIES = {
    "uniform3_r3": lambda s: gen.uniform_ksat(20000, 60000, 3, seed=500 + s),
    "tseitin": lambda s: gen.prepare(gen.tseitin(20000, 400, 40, seed=300 + s))[0],
    "heavytail": lambda s: gen.heavytail(20000, 80000, seed=400 + s),
}
OPTIONS = [{}, {"phases": 1, "bce": 1}, {"all": 1}]


@pytest.mark.parametrize("family", sorted(FAMILIES))
@pytest.mark.parametrize("opt", range(len(OPTIONS)))
def test_families(family, opt):
    cnf = FAMILIES[family](0)
    print("equisat family", family, OPTIONS[opt], flush=True)
    g = gpu(cnf.maxvar, cnf.offsets, cnf.lits, cnf.trail, **OPTIONS[opt])
    assert check(cnf.maxvar, cnf.offsets, cnf.lits, cnf.trail, None, g, offsets=True)
