"""The oracle against the reference CPU simplifier run live on seeded
random formulas (needs oracle/_ref/pfref, built from /root/reference by
oracle/build_ref.sh; skipped where it is absent).  REF mode must match
every output array bit for bit."""
import os
import subprocess
import tempfile

import pytest

from oracle import oracle
from parafrost_amd import gen, pfd
from tests.common import diff, oracle_kw
from tests.fuzz_cases import case

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "pfref")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference harness not built")


def ref_args(opts):
    out = []
    for k, v in opts.items():
        if k in ("phases", "mupos", "muneg"):
            out.append(f"--{k}={v}")
        else:
            out.append(f"-{k}" if v else f"-no-{k}")
    return out


@pytest.mark.parametrize("i", range(60))
def test_random_formulas(i):
    cnf, flags, lbd, ro, oo = case(i)
    if flags is not None:
        pytest.skip("the reference parser reads original clauses only")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd")] + ref_args(ro),
                       capture_output=True, timeout=120)
        if not os.path.exists(os.path.join(d, "i.pfd")):
            pytest.skip("refuted by the reference parser")
        inp, out = pfd.read(os.path.join(d, "i.pfd")), pfd.read(os.path.join(d, "o.pfd"))
    r = oracle.simplify_state(inp, det=0, **oo)
    d = diff(r, out, units_as_set=False)
    assert not d, d


@pytest.mark.parametrize("i", range(60))
def test_random_formulas_proof(i):
    """The oracle's DRAT lines equal the reference simplifier's (run with
    -proof; the lines written after rootify/shrinkTop)."""
    cnf, flags, lbd, ro, oo = case(i)
    if flags is not None:
        pytest.skip("the reference parser reads original clauses only")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        drat = os.path.join(d, "p.drat")
        subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd"), "-proof",
                        f"--proofout={drat}"] + ref_args(ro), capture_output=True, timeout=120)
        if not os.path.exists(os.path.join(d, "i.pfd")):
            pytest.skip("refuted by the reference parser")
        with open(os.path.join(d, "o.pfd.proofpos")) as f:
            pos = int(f.read())
        with open(drat, "rb") as f:
            ref_proof = f.read()[pos:]
        inp = pfd.read(os.path.join(d, "i.pfd"))
    r = oracle.simplify_state(inp, det=0, proof_en=1, **oo)
    assert r.proof == ref_proof


@pytest.mark.parametrize("i", range(0, 60, 5))
def test_aggressive_sort(i):
    """-aggresivesort: the output clauses in the stable order of the list key
    (the reference's newBeginning, src/cpu/simplify.cpp:263)."""
    cnf, flags, lbd, ro, oo = case(i)
    if flags is not None:
        pytest.skip("the reference parser reads original clauses only")
    ro = dict(ro, aggresivesort=1)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd")] + ref_args(ro),
                       capture_output=True, timeout=120)
        if not os.path.exists(os.path.join(d, "i.pfd")):
            pytest.skip("refuted by the reference parser")
        inp, out = pfd.read(os.path.join(d, "i.pfd")), pfd.read(os.path.join(d, "o.pfd"))
    r = oracle.simplify_state(inp, det=0, aggr_sort=1, **oo)
    d = diff(r, out, units_as_set=False)
    assert not d, d


def run_ref_learnt(cnf, ro, d, seed, pct=30, proof=False):
    """The reference with PFREF_LEARNT: a seeded share of the parsed clauses
    turned into learnt clauses (lbd 1..8, usage 0..2) before the simplify
    call, as the solver's inprocessing calls hand them over.  Returns the
    input and output states (None when the parser refutes the formula)."""
    p = os.path.join(d, "f.cnf")
    gen.write_dimacs(cnf, p)
    extra = ["-proof", f"--proofout={os.path.join(d, 'p.drat')}"] if proof else []
    env = dict(os.environ, PFREF_LEARNT=f"{seed},{pct}")
    subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd")] + extra + ref_args(ro),
                   capture_output=True, timeout=120, env=env)
    if not os.path.exists(os.path.join(d, "i.pfd")):
        return None, None
    return pfd.read(os.path.join(d, "i.pfd")), pfd.read(os.path.join(d, "o.pfd"))


@pytest.mark.parametrize("i", range(60))
def test_random_formulas_learnt(i):
    """Learnt clauses in the input (the inprocessing call): the oracle in REF
    mode equals the reference on every output array, learnt flags, lbd and
    usage included (subsumer promotion, src/cpu/subsume.h:113; usage and lbd
    of strengthened clauses, bumpShrunken; learnt resolvents,
    src/cpu/elimination.cpp:28-39)."""
    cnf, _, _, ro, oo = case(i)
    with tempfile.TemporaryDirectory() as d:
        inp, out = run_ref_learnt(cnf, ro, d, seed=i, pct=(20, 35, 50)[i % 3])
    if inp is None:
        pytest.skip("refuted by the reference parser")
    assert (inp.flags & 1).any() or inp.sizes.shape[0] < 4, "no learnt clause was planted"
    r = oracle.simplify_state(inp, det=0, **oo)
    d = diff(r, out, units_as_set=False)
    assert not d, d


@pytest.mark.parametrize("i", range(0, 60, 3))
def test_random_formulas_learnt_proof(i):
    """The DRAT lines with learnt clauses in the input equal the
    reference's."""
    cnf, _, _, ro, oo = case(i)
    with tempfile.TemporaryDirectory() as d:
        inp, out = run_ref_learnt(cnf, ro, d, seed=1000 + i, proof=True)
        if inp is None:
            pytest.skip("refuted by the reference parser")
        with open(os.path.join(d, "o.pfd.proofpos")) as f:
            pos = int(f.read())
        with open(os.path.join(d, "p.drat"), "rb") as f:
            ref_proof = f.read()[pos:]
    r = oracle.simplify_state(inp, det=0, proof_en=1, **oo)
    assert r.proof == ref_proof


@pytest.mark.parametrize("name,make", [
    ("u3", lambda: gen.uniform_ksat(3000, 12780, 3, seed=71)),
    ("ts", lambda: gen.tseitin(3000, seed=72)),
    ("ht", lambda: gen.heavytail(3000, 12000, seed=73)),
])
def test_structured_learnt(name, make):
    """Larger structured formulas (uniform 3-SAT, Tseitin, heavy-tailed)
    with 30 % learnt clauses, default options and the one-round BCE set."""
    cnf = make()
    for ro, oo in (({}, {}), ({"phases": 1, "bce": 1}, {"phases": 1, "bce_en": 1})):
        with tempfile.TemporaryDirectory() as d:
            inp, out = run_ref_learnt(cnf, ro, d, seed=7)
        if inp is None:
            pytest.skip("refuted by the reference parser")
        r = oracle.simplify_state(inp, det=0, **oo)
        d = diff(r, out, units_as_set=False)
        assert not d, (ro, d)
