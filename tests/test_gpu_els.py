"""Parity of the device ELS (pfs_decompose) with the reference's
Solver::decompose rounds.  Integer work: bit-exact on the clause order and
literal sets, the stack records (in order), the variable states and the
trail's set.

* the ELS fixtures uploaded from the reference's input, and the same
  formulas parsed from text on the device;
* random formulas with learnt clauses and assumed variables against the
  oracle; ELS followed by a simplify call against the oracle;
* a planted-equivalence formula with 2,000,000 variables: the
  representatives are the smallest literals of the strongly connected
  components computed by scipy, and every clause is its own image.
"""
import dataclasses
import json
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle import els, oracle
from parafrost_amd import gen, pfd
from parafrost_amd.simplifier import PfsError, Simplifier
from tests.common import GOLDEN, diff

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu
with open(os.path.join(GOLDEN, "els_manifest.json")) as f:
    EMANIFEST = json.load(f)


def lists(sizes, lits):
    off = np.zeros(sizes.shape[0] + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    return [lits[off[k]:off[k + 1]].tolist() for k in range(sizes.shape[0])]


def check(s, info, ref, trail0):
    assert info["status"] == ref.status
    if not ref.status:
        assert s.simplify() == 0
        return
    r = s.result(2)
    assert [sorted(c) for c in lists(r.sizes, r.lits)] == [sorted(c) for c in lists(ref.sizes, ref.lits)]
    assert r.resolved.tolist() == ref.resolved.tolist()
    assert np.array_equal(r.vstate, ref.vstate)
    assert sorted(list(trail0) + r.units.tolist()) == sorted(ref.units.tolist())
    assert sorted(s.trail().tolist()) == sorted(ref.units.tolist())


@pytest.mark.parametrize("entry", EMANIFEST, ids=[e["name"] for e in EMANIFEST])
def test_fixture_uploaded(entry):
    print("els fixture", entry["name"], flush=True)
    inp = pfd.read(os.path.join(GOLDEN, entry["input"]))
    ref = pfd.read(os.path.join(GOLDEN, entry["output"]))
    s = Simplifier(0)
    try:
        s.upload_state(inp)
        info = s.decompose(entry["rounds"] or 0)
        check(s, info, ref, inp.units.tolist())
    finally:
        s.close()


@pytest.mark.parametrize("entry", EMANIFEST, ids=[e["name"] for e in EMANIFEST])
def test_fixture_parsed(entry):
    cnf = {n: c for n, c, _ in make_golden.els_cases()}[entry["name"]]
    ref = pfd.read(os.path.join(GOLDEN, entry["output"]))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.cnf")
        gen.write_dimacs(cnf, p)
        with open(p, "rb") as f:
            text = f.read()
    s = Simplifier(0, decomposemin=entry["rounds"] or 2)
    try:
        s.parse_dimacs(text)
        trail0 = s.trail().tolist()
        info = s.decompose()
        check(s, info, ref, trail0)
    finally:
        s.close()


def oracle_els(cnf, learnt, assumed, rounds=2):
    cls = lists(np.diff(cnf.offsets).astype(np.uint32), cnf.lits)
    vstate = [0] * (cnf.maxvar + 1)
    return els.els(cnf.maxvar, cls, learnt, [], vstate, rounds=rounds, assumed=assumed)


@pytest.mark.parametrize("seed", range(6))
def test_learnts_and_assumed_match_oracle(seed):
    rng = np.random.default_rng(seed)
    cnf = gen.equivalences(3000, 6000, 200, 5, seed=100 + seed, anti=5 * (seed % 2))
    learnt = (rng.random(cnf.nclauses) < 0.3).tolist()
    assumed = set(int(v) for v in rng.choice(np.arange(1, cnf.maxvar + 1), 40, replace=False))
    st, o = oracle_els(cnf, learnt, assumed)
    flags = np.array([1 if x else 0 for x in learnt], np.uint8)
    lbd = np.where(flags == 1, 3, 0).astype(np.uint32)
    am = np.zeros(cnf.maxvar + 1, np.uint8)
    am[list(assumed)] = 1
    s = Simplifier(0)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits, flags=flags, lbd=lbd, assumed=am)
        info = s.decompose()
        assert info["status"] == st
        if not st:
            return
        r = s.result(2)
    finally:
        s.close()
    # the resident order is originals, then learnts (the upload's extraction order)
    keep = [(c, lr) for c, lr in zip(o.clauses, o.learnt) if c is not None]
    want = [sorted(c) for c, lr in keep if not lr] + [sorted(c) for c, lr in keep if lr]
    assert [sorted(c) for c in lists(r.sizes, r.lits)] == want
    assert r.resolved.tolist() == o.stack
    assert r.vstate.tolist() == o.vstate
    assert sorted(r.units.tolist()) == sorted(o.trail)


AFTER = [e for e in EMANIFEST if e["name"] in ("els_eq_b", "els_eq_units_b", "els_eq_2sat", "els_ts_g2000")]


@pytest.mark.parametrize("entry", AFTER, ids=[e["name"] for e in AFTER])
def test_simplify_after_els_matches_oracle(entry):
    """A simplify call on the formula ELS left equals the oracle run on the
    reference's ELS output."""
    inp = pfd.read(os.path.join(GOLDEN, entry["input"]))
    ref = pfd.read(os.path.join(GOLDEN, entry["output"]))
    s = Simplifier(0, phases=1, bce=1)
    try:
        s.upload_state(inp)
        s.decompose(entry["rounds"] or 0)
        mid = s.result(2)
        g = s.result(s.simplify())
    finally:
        s.close()
    g = dataclasses.replace(g, units=g.units[mid.units.shape[0]:], resolved=g.resolved[mid.resolved.shape[0]:])
    off = np.zeros(ref.sizes.shape[0] + 1, np.uint64)
    np.cumsum(ref.sizes, out=off[1:])
    r = oracle.simplify(ref.maxvar, off, ref.lits, ref.flags, ref.lbd, ref.vstate, trail=ref.units, det=1, phases=1,
                        bce_en=1)
    d = diff(g, r)
    assert not d, d


def test_full_size_planted_equivalences():
    """2,000,000 variables, 6,000,000 ternary clauses, 200,000 classes of 5
    literals: the substitution equals the map to the smallest literal of
    each strongly connected component (scipy), clause by clause."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    cnf = gen.equivalences(2_000_000, 6_000_000, 200_000, 5, seed=0x5EED0007, chords=0.4)
    s = Simplifier(0)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        info = s.decompose(1)
        r = s.result(2)
    finally:
        s.close()
    sizes = np.diff(cnf.offsets).astype(np.int64)
    bins = np.nonzero(sizes == 2)[0]
    a = cnf.lits[cnf.offsets[bins].astype(np.int64)].astype(np.int64)
    b = cnf.lits[cnf.offsets[bins].astype(np.int64) + 1].astype(np.int64)
    n2 = 2 * cnf.maxvar + 2
    g = csr_matrix((np.ones(2 * bins.shape[0], np.int8), (np.concatenate([a ^ 1, b ^ 1]), np.concatenate([b, a]))),
                   shape=(n2, n2))
    _, comp = connected_components(g, directed=True, connection="strong")
    rep_of = np.full(comp.max() + 1, n2, np.int64)
    np.minimum.at(rep_of, comp, np.arange(n2))
    rep = rep_of[comp]
    assert info["status"] == 2
    subst = np.nonzero(rep[2 * np.arange(1, cnf.maxvar + 1)] != 2 * np.arange(1, cnf.maxvar + 1))[0] + 1
    assert info["substituted"] == subst.shape[0]
    assert np.array_equal(np.nonzero(r.vstate == 4)[0], subst)
    # the stack: one binary record per substituted variable, in variable order
    recs = r.resolved.reshape(-1, 5)
    assert np.array_equal(recs[:, 0], 2 * subst)
    assert np.array_equal(recs[:, 1], rep[2 * subst] ^ 1)
    # every input clause maps to its image: sorted, duplicates dropped,
    # tautologies (the binary cycle clauses) removed
    out = []
    for k in range(cnf.nclauses):
        c = sorted(set(int(x) for x in rep[cnf.lits[int(cnf.offsets[k]):int(cnf.offsets[k + 1])]]))
        if any((x ^ 1) in c for x in c):
            continue
        out.append(c)
        if k > 300_000:
            break
    got = lists(r.sizes, r.lits)
    assert got[:len(out)] == out


def test_decompose_errors():
    cnf = gen.equivalences(200, 500, 10, 4, seed=1)
    s = Simplifier(0)
    try:
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.decompose()
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        s.append(cnf.maxvar + 1, np.array([0, 1], np.uint64), np.array([2 * cnf.maxvar + 2], np.uint32))
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.decompose()
    finally:
        s.close()
