"""The GPU build's election limits in the oracle (electionsmax and
lcveclausemax, src/gpu/lcve.cu:43-52 and 94; the CPU simplifier has neither,
so the compiled reference cannot pin them).  Checked by what the rules fix:

* a limit at or above every occurrence count / clause length changes nothing;
* electionsmax below every candidate's largest literal count stops the
  election at its first candidate: nothing is elected, the formula comes back
  unchanged (LCVE fails lcvemin, src/cpu/simplify.cpp:184);
* lcveclausemax 1 passes over every candidate (every clause has two or more
  literals): nothing is elected either;
* electionsmax cuts the election's prefix, so it elects no more than the
  unlimited election in the first round (and fewer where it bites).
"""
import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests.common import diff


def _formulas():
    yield gen.uniform_ksat(400, 1500, 3, seed=11)
    yield gen.heavytail(3000, 12000, seed=12)
    yield gen.prepare(gen.tseitin(3000, 100, 20, seed=13))[0]


def _run(cnf, **kw):
    return oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1,
                           phases=1, bce_en=1, **kw)


def _max_count(cnf):
    c = np.bincount(cnf.lits, minlength=2 * cnf.maxvar + 2)
    return int(c.max())


@pytest.mark.parametrize("k", range(3))
def test_limits_above_everything_change_nothing(k):
    cnf = list(_formulas())[k]
    base = _run(cnf)
    lmax = int(np.diff(cnf.offsets).max())
    assert not diff(_run(cnf, lcve_max_occurs=_max_count(cnf)), base)
    assert not diff(_run(cnf, lcve_clause_max=lmax), base)
    assert list(_run(cnf, lcve_max_occurs=_max_count(cnf)).elected) == list(base.elected)


@pytest.mark.parametrize("k", range(3))
def test_limits_that_stop_everything(k):
    cnf = list(_formulas())[k]
    sizes = np.diff(cnf.offsets).astype(np.int64)
    for kw in ({"lcve_clause_max": 1},):
        r = _run(cnf, **kw)
        assert list(r.elected) == [0]
        assert r.status == 2 and np.array_equal(r.sizes, sizes)
    # every variable of these formulas has a literal with two or more occurrences
    r = _run(cnf, lcve_max_occurs=1)
    counts = np.bincount(cnf.lits, minlength=2 * cnf.maxvar + 2).reshape(-1, 2)[1:]
    if (counts.max(axis=1)[counts.sum(axis=1) > 0] > 1).all():
        assert list(r.elected) == [0]


@pytest.mark.parametrize("k", range(3))
def test_limits_elect_fewer(k):
    cnf = list(_formulas())[k]
    base = int(_run(cnf).elected[0])
    cut = int(_run(cnf, lcve_max_occurs=8).elected[0])
    assert cut < base
    assert int(_run(cnf, lcve_max_occurs=20).elected[0]) <= base
