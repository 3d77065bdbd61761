"""The oracle at full size: its deterministic-mode output on the 2M-gate
Tseitin circuit (6.77M clauses, the benchmark's one-round options) equals
the digest kept in tests/golden/fullsize.json (which the GPU tests compare
against), and the reference CPU simplifier (oracle/_ref/pfref, where built)
on the same formula equals the oracle in REF mode, proof lines included.
The other full-size legs (the Tseitin circuit under default options, the
10M-variable heavy-tailed formula) take minutes on one core: they run with
PFS_SLOW_TESTS=1 (tests/golden/make_fullsize.py recorded their agreement)."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import gen, pfd
from tests.common import GOLDEN, oracle_kw, result_digest

with open(os.path.join(GOLDEN, "fullsize.json")) as f:
    FULL = {e["name"]: e for e in json.load(f)}
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "pfref")
SLOW = os.environ.get("PFS_SLOW_TESTS") == "1"
_CACHE = {}


def workload(name):
    if name not in _CACHE:
        _CACHE.clear()
        _CACHE[name] = gen.workload(name)
    return _CACHE[name]


@pytest.mark.parametrize("name", sorted(FULL))
def test_oracle_matches_fullsize_digest(name):
    e = FULL[name]
    if name != "tseitin.p1bce" and not SLOW:
        pytest.skip("minutes on one core: PFS_SLOW_TESTS=1")
    cnf = workload(e["workload"])
    vs = np.zeros(cnf.maxvar + 1, np.uint8)
    vs[np.asarray(cnf.trail, np.int64) >> 1] = 2
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1, vstate=vs,
                        proof_en=1, **oracle_kw(e["options"]))
    assert result_digest(r) == e["oracle_det"]
    if "reference" in e:
        assert e["reference"]["oracle_ref_equal"] and e["reference"]["input_is_prepared_formula"]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference harness not built")
def test_reference_matches_oracle_full_tseitin():
    """pfref on the full Tseitin text (its own parse, root propagation and
    one-round simplify with proof) equals the oracle in REF mode on the
    reference's input state, and the DET-mode digest."""
    cnf = workload("tseitin")
    units = np.asarray(cnf.trail, np.uint32)
    full = gen.Cnf(cnf.maxvar, np.concatenate([cnf.offsets, cnf.offsets[-1] + 1 + np.arange(units.shape[0],
                                                                                          dtype=np.uint64)]),
                   np.concatenate([cnf.lits, units]))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.cnf")
        gen.write_dimacs(full, p)
        drat = os.path.join(d, "p.drat")
        subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd"), "--phases=1", "-bce", "-proof",
                        f"--proofout={drat}"], capture_output=True, timeout=900, check=True)
        inp, out = pfd.read(os.path.join(d, "i.pfd")), pfd.read(os.path.join(d, "o.pfd"))
        with open(os.path.join(d, "o.pfd.proofpos")) as f:
            pos = int(f.read())
        with open(drat, "rb") as f:
            out.proof = f.read()[pos:]
    r = oracle.simplify_state(inp, det=0, proof_en=1, phases=1, bce_en=1)
    assert result_digest(r) == result_digest(out)
    assert result_digest(out) == FULL["tseitin.p1bce"]["oracle_det"]
