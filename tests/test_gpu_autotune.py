"""The launch-shape tuner (option "autotune", pfs_launch_shape): on a
formula class's first call with at least 2^20 clauses the occurrence-table
tile and the election's lanes per candidate and LDS reservation are chosen
by timing each candidate; later calls of the class reuse the choice; shapes set by option
are kept.  The output is the same whatever the shapes."""
import numpy as np
import pytest

from parafrost_amd import gen
from tests.common import diff

pytestmark = pytest.mark.gpu
TILES = (11, 12, 13, 14)
LDS = (0, 10240, 20480)
LANES = (1, 2, 4, 8)


@pytest.fixture(scope="module")
def formula():
    cnf = gen.heavytail(310_000, 1_234_567, seed=977)
    assert cnf.nclauses >= 1 << 20
    return cnf


def _run(cnf, **opts):
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0, phases=1, bce=1, **opts)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
        r = s.result(s.simplify())
        return r, s.launch_shape()
    finally:
        s.close()


def test_tuned_once_per_class_and_same_output(formula):
    a, sa = _run(formula)
    b, sb = _run(formula)
    c, sc = _run(formula, autotune=0, otptile=12, mislds=10240, mislanes=4)
    print("shapes", sa, sb, sc, flush=True)
    # the first call tuned the class unless another context of the class
    # already had; the second reuses the choice
    key = ("otp_tile_log2", "mis_lds", "mis_lanes")
    assert sa["otp_tile_log2"] in TILES and sa["mis_lds"] in LDS and sa["mis_lanes"] in LANES
    assert not sb["tuned"] and [sb[k] for k in key] == [sa[k] for k in key]
    assert not sc["tuned"] and [sc[k] for k in key] == [12, 10240, 4]
    for other in (b, c):
        d = diff(a, other, units_as_set=True)
        assert not d, d


def test_small_formulas_are_not_tuned():
    cnf = gen.uniform_ksat(20000, 85200, 3, seed=5)
    _, sh = _run(cnf)
    assert not sh["tuned"] and sh["mis_lds"] == 10240 and sh["mis_lanes"] == 1
