"""The variable map restatement (oracle/vmap.py) against the reference.

* golden: for every fixture whose run ends unsolved, the restatement
  applied to the reference's simplifier output reproduces the state that
  the reference's own Solver::map(true) / newBeginning() left (PFM1 dumps
  written by oracle/ref_harness.cpp): applied flag, maxVar, vorg, trail,
  model literal table, variable states, clause arena;
* closed-form properties of the map on synthetic states;
* planted mistakes in the restatement are caught by the golden dumps.
"""
import os

import numpy as np
import pytest

from oracle import vmap
from tests.common import GOLDEN, load_manifest, read_golden

MAPPED = [e for e in load_manifest() if "map" in e]


def oracle_map(entry, **kw):
    inp, out = read_golden(entry)
    mv = out.maxvar
    trail = np.concatenate([inp.units, out.units]).astype(np.uint32)
    assign = np.zeros(mv + 1, np.uint32)
    assign[trail >> 1] = trail + 1
    return vmap.map_state(mv, out.vstate, assign, out.sizes, out.flags, out.lbd, out.lits, trail=trail, **kw)


def mismatches(got, ref):
    bad = []
    for k in ref:
        a, b = np.asarray(got[k]), np.asarray(ref[k])
        if a.shape != b.shape or not np.array_equal(a, b):
            bad.append(k)
    return bad


def test_golden_map_fixtures_present():
    assert len(MAPPED) >= 30
    applied = [vmap.read_pfm(os.path.join(GOLDEN, e["map"]))["applied"] for e in MAPPED]
    # both branches of the write-back are covered
    assert 0 < sum(applied) < len(applied)


@pytest.mark.parametrize("entry", MAPPED, ids=[e["name"] for e in MAPPED])
def test_map_matches_reference(entry):
    ref = vmap.read_pfm(os.path.join(GOLDEN, entry["map"]))
    got = oracle_map(entry)
    assert not mismatches(got, ref), mismatches(got, ref)


def test_planted_mistakes_are_caught():
    """Each plausible slip in the restatement fails at least one fixture."""
    refs = [(e, vmap.read_pfm(os.path.join(GOLDEN, e["map"]))) for e in MAPPED]
    applied = [(e, r) for e, r in refs if r["applied"]]
    real_initiate, real_map_lits = vmap.initiate, vmap.map_lits

    def flipped_first(maxvar, vstate, assign):  # wrong polarity of the stand-in literal
        m, f, t = real_initiate(maxvar, vstate, assign)
        return m, f, (t ^ 1) if f else t

    def no_first(maxvar, vstate, assign):  # the frozen stand-in not kept
        m, f, t = real_initiate(maxvar, vstate, assign)
        if not f:
            return m, f, t
        m = m.astype(np.int64)
        m[f] = 0
        m[f + 1:] = np.where(m[f + 1:] > 0, m[f + 1:] - 1, 0)
        return m.astype(np.uint32), 0, 0

    def last_first(maxvar, vstate, assign):  # the last frozen variable instead of the first
        v = np.arange(maxvar + 1)
        cand = np.nonzero((v > 0) & (vstate[: maxvar + 1] != 0) & (assign[: maxvar + 1] != 0))[0]
        if cand.size < 2:
            return real_initiate(maxvar, vstate, assign)
        f = int(cand[-1])
        keep = (v > 0) & ((vstate[: maxvar + 1] == 0) | (v == f))
        m = np.where(keep, np.cumsum(keep), 0).astype(np.uint32)
        return m, f, 2 * int(m[f]) + ((int(assign[f]) - 1) & 1)

    def sign_lost(lits, mapped, first, first_true, assign):  # sign dropped on direct mapping
        out = real_map_lits(lits, mapped, first, first_true, assign)
        return np.where((np.asarray(lits) & 1) == 1, out & ~np.uint32(1), out).astype(np.uint32)

    plants = [("initiate", flipped_first), ("initiate", no_first), ("initiate", last_first),
              ("map_lits", sign_lost)]
    try:
        for name, fn in plants:
            setattr(vmap, name, fn)
            caught = any(mismatches(oracle_map(e), r) for e, r in applied)
            setattr(vmap, name, real_initiate if name == "initiate" else real_map_lits)
            assert caught, f"planted mistake {fn.__name__} not caught"
        # map when the threshold says no (canMap's strict comparison)
        unapplied = [(e, r) for e, r in refs if not r["applied"]]
        assert any(mismatches(oracle_map(e, force=True), r) for e, r in unapplied)
    finally:
        vmap.initiate, vmap.map_lits = real_initiate, real_map_lits


def synthetic(seed, maxvar=200):
    rng = np.random.default_rng(seed)
    vstate = np.zeros(maxvar + 1, np.uint8)
    vstate[1:] = rng.choice([0, 0, 1, 2, 4], maxvar)
    assign = np.zeros(maxvar + 1, np.uint32)
    frozen = np.nonzero(vstate == 2)[0]
    assign[frozen] = 2 * frozen + rng.integers(0, 2, frozen.shape[0]) + 1
    return vstate, assign


@pytest.mark.parametrize("seed", range(8))
def test_map_closed_form(seed):
    maxvar = 200
    vstate, assign = synthetic(seed, maxvar)
    mapped, first, first_true = vmap.initiate(maxvar, vstate, assign)
    frozen = np.nonzero((vstate != 0) & (assign != 0))[0]
    assert first == (int(frozen[0]) if frozen.size else 0)
    kept = np.nonzero(mapped)[0]
    # dense, order preserving: the kept variables become 1..n in order
    assert np.array_equal(mapped[kept], np.arange(1, kept.size + 1))
    assert set(kept.tolist()) == set(np.nonzero(vstate == 0)[0][1:].tolist()) | ({first} if first else set())
    # every literal of an active variable keeps its sign; an assigned
    # literal maps to the stand-in with its value; others vanish
    lits = np.arange(2, 2 * maxvar + 2, dtype=np.uint32)
    m = vmap.map_lits(lits, mapped, first, first_true, assign)
    for lit, x in zip(lits.tolist(), m.tolist()):
        v = lit >> 1
        if vstate[v] == 0:
            assert x == 2 * mapped[v] + (lit & 1)
        elif assign[v]:
            true = lit == assign[v] - 1
            assert x == (first_true if true else first_true ^ 1)
        else:
            assert x == 0
    assert vmap.map_lits(np.zeros(3, np.uint32), mapped, first, first_true, assign).tolist() == [0, 0, 0]


def test_can_map_is_strict():
    vstate = np.zeros(11, np.uint8)
    vstate[1:3] = 1  # 2 of 10 inactive
    assert not vmap.can_map(vstate, 10, 0.2)
    vstate[3] = 2
    assert vmap.can_map(vstate, 10, 0.2)
    assert not vmap.can_map(vstate, 10, 1.0)
