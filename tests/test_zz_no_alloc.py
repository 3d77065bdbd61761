"""pfs_map_lits_noalloc (the entry point that allocates nothing) against
pfs_map_lits through the C ABI: equal outputs on two calls in a row, the
second with a smaller input, on a formula whose variable map renumbers and
drops variables; sizes beyond the fixed scratch are refused."""
import numpy as np
import pytest

from parafrost_amd import gen

pytestmark = pytest.mark.gpu
SCRATCH_WORDS = 1 << 20


def _mapped_context():
    from parafrost_amd.simplifier import Simplifier
    cnf = gen.prepare(gen.tseitin(20000, 400, 40, seed=31))[0]
    s = Simplifier(0, phases=1, bce=1)
    s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
    s.simplify()
    info = s.map(force=True)
    assert info["applied"] and info["new_maxvar"] < info["old_maxvar"]
    return s, info["old_maxvar"]


def test_noalloc_equals_map_lits():
    s, nv = _mapped_context()
    try:
        rng = np.random.default_rng(5)
        for n in (SCRATCH_WORDS, 777):  # full scratch, then a smaller call over stale contents
            lits = (2 * rng.integers(1, nv + 1, n) + rng.integers(0, 2, n)).astype(np.uint32)
            lits[rng.integers(0, n, max(1, n // 50))] = 0  # zeros stay zeros
            want = s.map_lits(lits)
            got = s.map_lits_noalloc(lits)
            assert np.array_equal(got, want), n
            assert not np.array_equal(want, lits)  # the map changed them
    finally:
        s.close()


def test_noalloc_refuses_beyond_scratch():
    from parafrost_amd.simplifier import PfsError
    s, nv = _mapped_context()
    try:
        lits = np.full(SCRATCH_WORDS + 1, 2, np.uint32)
        with pytest.raises(PfsError, match="BAD_INPUT.*at most"):
            s.map_lits_noalloc(lits)
    finally:
        s.close()
