"""Batched simplification (pfs_batch_simplify): many resident formulas
simplified by concurrent native workers, one stream per context.  Every
formula's output must equal its plain single-formula run bit for bit, and
the plain runs equal the oracle."""
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests.common import diff

pytestmark = pytest.mark.gpu


def _contexts(cnfs, **opts):
    from parafrost_amd.simplifier import Simplifier
    out = []
    for c in cnfs:
        s = Simplifier(0, keep_input=True, proof=1, **opts)
        s.upload(c.maxvar, c.offsets, c.lits, trail=c.trail)
        out.append(s)
    return out


@pytest.mark.parametrize("workers", [1, 4, 16])
def test_batch_equals_plain_runs(workers):
    from parafrost_amd.simplifier import batch_simplify
    cnfs = gen.batch(24, seed=0xBA7C0 + workers, cmax=3e4)
    print("batch", workers, sum(c.nclauses for c in cnfs), flush=True)
    ctx = _contexts(cnfs, phases=2, bce=1)
    try:
        plain = []
        for s in ctx:
            plain.append(s.result(s.simplify()))
        for s in ctx:
            s.rewind()
        res, ms = batch_simplify(ctx, workers)
        assert len(ms) == len(ctx)
        for i, s in enumerate(ctx):
            b = s.result(res[i])
            d = diff(b, plain[i], units_as_set=False)
            assert not d, (i, d)
            assert b.proof == plain[i].proof
    finally:
        for s in ctx:
            s.close()


def test_batch_matches_oracle():
    from parafrost_amd.simplifier import batch_simplify
    cnfs = gen.batch(12, seed=0xBA7C9, cmax=1e4)
    ctx = _contexts(cnfs)
    try:
        res, _ = batch_simplify(ctx, 6)
        for i, (c, s) in enumerate(zip(cnfs, ctx)):
            import numpy as np
            vs = np.zeros(c.maxvar + 1, np.uint8)
            vs[c.trail >> 1] = 2
            r = oracle.simplify(c.maxvar, c.offsets, c.lits, nassigned=int(c.trail.shape[0]), vstate=vs, det=1)
            d = diff(s.result(res[i]), r)
            assert not d, (i, d)
    finally:
        for s in ctx:
            s.close()
