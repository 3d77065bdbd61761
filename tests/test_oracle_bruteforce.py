"""Pins the oracle by what the mathematics fixes, on formulas small enough to
enumerate (<= 16 variables):

* a refutation is only reported for unsatisfiable formulas;
* the simplified formula plus the new units is satisfiable iff the input is
  (BVE, BCE and the gate substitutions preserve satisfiability);
* every model of the simplified formula, extended over the eliminated-variable
  stack, satisfies the input formula (model reconstruction);
* with elimination off (SUB and ERE only), the output is logically
  equivalent to the input over all 2^n assignments.

A plausible slip in the restatement (a dropped resolvent, a wrong witness,
a wrong literal in a strengthened clause) breaks one of these.
"""
import itertools
import random

import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests.common import extend_model


def clauses_of(offsets, lits):
    return [[int(x) for x in lits[offsets[i]:offsets[i + 1]]] for i in range(len(offsets) - 1)]


def sat_table(nv, clauses):
    """Boolean vector over all 2^nv assignments (bit v-1 = value of var v)."""
    a = np.arange(1 << nv, dtype=np.int64)
    ok = np.ones(1 << nv, bool)
    for c in clauses:
        s = np.zeros(1 << nv, bool)
        for l in c:
            bit = (a >> ((l >> 1) - 1)) & 1
            s |= (bit == (0 if (l & 1) else 1))
        ok &= s
    return ok


def random_formula(rng, nv):
    m = rng.randint(nv, 5 * nv)
    cls = []
    for _ in range(m):
        k = rng.choice([2, 2, 3, 3, 4])
        vs = rng.sample(range(1, nv + 1), k)
        cls.append([v if rng.random() < 0.5 else -v for v in vs])
        if rng.random() < 0.1:
            cls.append(list(cls[-1]))
    return gen.from_clauses(nv, cls)


OPTS = [dict(), dict(phases=1, bce_en=1), dict(all_en=1), dict(all_en=1, mu_pos=10, mu_neg=10, phases=4)]


@pytest.mark.parametrize("seed", range(60))
def test_equisatisfiable_and_models_extend(seed):
    rng = random.Random(seed)
    nv = rng.randint(4, 14)
    cnf = random_formula(rng, nv)
    orig = clauses_of(cnf.offsets, cnf.lits)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, **OPTS[seed % len(OPTS)])
    orig_sat = sat_table(nv, orig)
    if r.status == 0:
        assert not orig_sat.any(), "refutation reported for a satisfiable formula"
        return
    off = np.zeros(r.sizes.shape[0] + 1, np.uint64)
    np.cumsum(r.sizes, out=off[1:])
    simp = clauses_of(off, r.lits) + [[int(u)] for u in r.units]
    simp_sat = sat_table(nv, simp)
    assert orig_sat.any() == simp_sat.any()
    eliminated = {v for v in range(1, nv + 1) if r.vstate[v] == 1}
    # variables of the simplified formula must not be eliminated ones
    assert not any((l >> 1) in eliminated for c in simp for l in c)
    for m in np.nonzero(simp_sat)[0][:64]:
        values = {v: bool((m >> (v - 1)) & 1) for v in range(1, nv + 1)}
        for v in eliminated:
            values[v] = False
        extend_model(values, r.resolved)
        assert all(any(values[l >> 1] == (not (l & 1)) for l in c) for c in orig), "extended model fails"


@pytest.mark.parametrize("seed", range(30))
def test_sub_ere_preserve_equivalence(seed):
    rng = random.Random(1000 + seed)
    nv = rng.randint(4, 12)
    cnf = random_formula(rng, nv)
    orig = clauses_of(cnf.offsets, cnf.lits)
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, ve_en=0, ve_plus_en=0, phases=1, ere_en=1)
    if r.status == 0:
        assert not sat_table(nv, orig).any()
        return
    off = np.zeros(r.sizes.shape[0] + 1, np.uint64)
    np.cumsum(r.sizes, out=off[1:])
    simp = clauses_of(off, r.lits) + [[int(u)] for u in r.units]
    assert np.array_equal(sat_table(nv, orig), sat_table(nv, simp))
    assert r.resolved.shape[0] == 0


def test_bce_only_removes_blocked_clauses():
    # (x1 v x2) is blocked on x1 when every clause with -x1 contains -x2
    cnf = gen.from_clauses(3, [[1, 2], [-1, -2, 3], [-1, -2, -3], [2, 3], [-2, -3]])
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, bce_en=1, ve_en=0, ve_plus_en=0, sub_en=0,
                        ere_en=0, phases=1)
    assert r.status in (1, 2)
    orig = clauses_of(cnf.offsets, cnf.lits)
    off = np.zeros(r.sizes.shape[0] + 1, np.uint64)
    np.cumsum(r.sizes, out=off[1:])
    assert sat_table(3, orig).any() == sat_table(3, clauses_of(off, r.lits)).any()


def test_and_gate_substitution_counts():
    # x3 = x1 & x2 plus two uses of x3: substituting the gate yields the
    # gate-by-non-gate resolvents only
    cnf = gen.from_clauses(5, [[-3, 1], [-3, 2], [3, -1, -2], [3, 4], [-3, 5, 4]])
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, phases=1)
    assert r.status in (1, 2)
    orig = clauses_of(cnf.offsets, cnf.lits)
    off = np.zeros(r.sizes.shape[0] + 1, np.uint64)
    np.cumsum(r.sizes, out=off[1:])
    simp = clauses_of(off, r.lits) + [[int(u)] for u in r.units]
    assert sat_table(5, orig).any() == sat_table(5, simp).any()
