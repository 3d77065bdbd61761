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
