"""Parity of the HIP simplifier with the oracle (and through it with the
reference CPU simplifier).  Integer/Boolean work: bit-exact.

* golden fixtures: the GPU output equals the reference's own output;
* seeded formulas of every workload family at sizes the oracle finishes in
  seconds, under several option sets, equal the oracle's DET mode;
* the full-size uniform 3-SAT workload (20,000 variables, 85,200 clauses).
"""
import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests.common import diff, load_manifest, oracle_kw, read_golden

pytestmark = pytest.mark.gpu
MANIFEST = load_manifest()


def run_gpu(maxvar, offsets, lits, trail=None, flags=None, lbd=None, vstate=None, **opts):
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0, **opts)
    try:
        s.upload(maxvar, offsets, lits, flags, lbd, vstate, None, None, trail)
        return s.result(s.simplify())
    finally:
        s.close()


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
def test_golden(entry):
    from parafrost_amd.simplifier import simplify_state
    inp, out = read_golden(entry)
    print("golden", entry["name"], flush=True)
    g = simplify_state(inp, **entry["options"])
    d = diff(g, out)
    assert not d, d


FAMILIES = {
    "uniform3": lambda s: gen.uniform_ksat(3000, 12780, 3, seed=100 + s),
    "uniform2": lambda s: gen.uniform_ksat(2000, 1500, 2, seed=200 + s),
    "tseitin": lambda s: gen.prepare(gen.tseitin(20000, 400, 40, seed=300 + s))[0],
    "heavytail": lambda s: gen.heavytail(20000, 80000, seed=400 + s),
}
OPTIONS = [{}, {"phases": 1, "bce": 1}, {"all": 1}, {"all": 1, "mupos": 10, "muneg": 10, "phases": 3}]


@pytest.mark.parametrize("family", sorted(FAMILIES))
@pytest.mark.parametrize("seed", range(2))
@pytest.mark.parametrize("opt", range(len(OPTIONS)))
def test_family_matches_oracle(family, seed, opt):
    cnf = FAMILIES[family](seed)
    o = OPTIONS[opt]
    print("family", family, seed, o, cnf.nclauses, flush=True)
    g = run_gpu(cnf.maxvar, cnf.offsets, cnf.lits, cnf.trail, **o)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1,
                        vstate=_vstate(cnf), **oracle_kw(o))
    d = diff(g, r)
    assert not d, d


def _vstate(cnf):
    vs = np.zeros(cnf.maxvar + 1, np.uint8)
    vs[cnf.trail >> 1] = 2
    return vs


def test_learnt_clauses_match_oracle():
    rng = np.random.default_rng(5)
    cnf = gen.uniform_ksat(3000, 12000, 3, seed=77)
    flags = (rng.random(cnf.nclauses) < 0.3).astype(np.uint8) | (rng.integers(0, 3, cnf.nclauses) << 4).astype(np.uint8)
    lbd = rng.integers(1, 9, cnf.nclauses).astype(np.uint32)
    g = run_gpu(cnf.maxvar, cnf.offsets, cnf.lits, None, flags, lbd, all=1)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, flags, lbd, det=1, all_en=1)
    d = diff(g, r)
    assert not d, d


def test_full_uniform3_workload():
    cnf = gen.workload("uniform3")
    for o in ({"phases": 1, "bce": 1}, {}):
        g = run_gpu(cnf.maxvar, cnf.offsets, cnf.lits, **o)
        r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, det=1, **oracle_kw(o))
        d = diff(g, r)
        assert not d, (o, d)


def test_rewind_repeats_bit_exact():
    from parafrost_amd.simplifier import Simplifier
    cnf = gen.uniform_ksat(4000, 17040, 3, seed=9)
    s = Simplifier(0, keep_input=True, phases=1, bce=1)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        a = s.result(s.simplify())
        for _ in range(2):
            s.rewind()
            b = s.result(s.simplify())
            assert not diff(a, b, units_as_set=False)
    finally:
        s.close()


def test_bad_input_is_rejected():
    from parafrost_amd.simplifier import PfsError, Simplifier
    s = Simplifier(0)
    try:
        with pytest.raises(PfsError):
            s.upload(3, np.array([0, 2], np.uint64), np.array([2, 9], np.uint32))  # var 4 > maxvar
        with pytest.raises(PfsError):
            s.set_option("nosuchoption", 1)
        with pytest.raises(PfsError):
            s.simplify()  # nothing uploaded
    finally:
        s.close()


def test_memory_budget_admission():
    from parafrost_amd.simplifier import PfsError, Simplifier
    cnf = gen.uniform_ksat(2000, 8520, 3, seed=3)
    s = Simplifier(0, mem_budget=1 << 16)
    try:
        with pytest.raises(PfsError, match="ALLOC_FAIL"):
            s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
    finally:
        s.close()


from tests.fuzz_cases import case as fuzz_case  # noqa: E402


@pytest.mark.parametrize("i", range(150))
def test_fuzz_dense_formulas(i):
    """Small dense formulas: units from strengthening (re-run path),
    cross-variable ERE deletions, refutations, learnt clauses."""
    cnf, flags, lbd, ro, oo = fuzz_case(i)
    prep, refuted = (cnf, False) if flags is not None else gen.prepare(cnf)
    if refuted:
        pytest.skip("refuted at the root")
    g = run_gpu(prep.maxvar, prep.offsets, prep.lits, prep.trail, flags, lbd, **ro)
    r = oracle.simplify(prep.maxvar, prep.offsets, prep.lits, flags, lbd, vstate=_vstate(prep),
                        nassigned=int(prep.trail.shape[0]), det=1, **oo)
    d = diff(g, r)
    assert not d, (i, d)
