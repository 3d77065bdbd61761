"""The oracle against the reference CPU simplifier's own outputs.

The golden fixtures were written by the reference CPU simplifier compiled
from its sources (tests/golden/make_golden.py).  In REF mode the oracle
reproduces the reference's occurrence-list sort exactly, so every output
array must match bit for bit; in DET mode (the order the GPU implements)
the outputs must match as well on these fixtures.
"""
import numpy as np
import pytest

from oracle import oracle
from tests.common import diff, load_manifest, oracle_kw, read_golden

MANIFEST = load_manifest()


@pytest.mark.parametrize("entry", MANIFEST, ids=[e["name"] for e in MANIFEST])
@pytest.mark.parametrize("det", [0, 1], ids=["ref", "det"])
def test_oracle_matches_reference(entry, det):
    inp, out = read_golden(entry)
    r = oracle.simplify_state(inp, det=det, **oracle_kw(entry["options"]))
    d = diff(r, out, units_as_set=False)
    assert not d, d


def test_fixture_coverage():
    """The fixtures exercise eliminations, resolvents, units and refutations."""
    kinds = set()
    for e in MANIFEST:
        _, out = read_golden(e)
        kinds.add("unsat" if out.status == 0 else "sat" if out.status == 1 else "open")
        if out.status:
            if out.resolved.shape[0]:
                kinds.add("stack")
            if (out.flags & 2).any():
                kinds.add("resolvents")
            if out.units.shape[0]:
                kinds.add("units")
    assert {"stack", "resolvents", "units"} <= kinds, kinds
