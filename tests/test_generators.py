import numpy as np

from parafrost_amd import gen


def test_uniform_is_deterministic_and_well_formed():
    a = gen.uniform_ksat(500, 2130, 3, seed=5)
    b = gen.uniform_ksat(500, 2130, 3, seed=5)
    assert np.array_equal(a.lits, b.lits)
    assert a.nclauses == 2130 and a.nlits == 3 * 2130
    for i in range(0, a.nclauses, 97):
        c = a.clause(i) >> 1
        assert len(set(c.tolist())) == 3 and c.min() >= 1 and c.max() <= 500


def test_tseitin_shape():
    c = gen.tseitin(1000, 50, 10, seed=3)
    sizes = np.diff(c.offsets)
    assert c.maxvar == 1050
    assert (sizes[-10:] == 1).all() and set(np.unique(sizes[:-10]).tolist()) <= {2, 3}
    p, refuted = gen.prepare(c)
    if not refuted:
        assert p.trail.shape[0] >= 10
        assert (np.diff(p.offsets) >= 2).all()


def test_heavytail_lengths_and_skew():
    c = gen.heavytail(20000, 80000, seed=9)
    sizes = np.diff(c.offsets)
    assert sizes.min() >= 2 and sizes.max() <= 30
    assert 2.8 < sizes.mean() < 3.2
    occ = np.bincount(c.lits >> 1, minlength=20001)[1:]
    assert occ.max() > 50 * np.median(occ[occ > 0])


def test_prepare_propagates_units():
    c = gen.from_clauses(4, [[1], [-1, 2], [-2, 3, 4], [-3, -4]])
    p, refuted = gen.prepare(c)
    assert not refuted
    assert sorted(p.trail.tolist()) == [2, 4]
    assert p.nclauses == 2
