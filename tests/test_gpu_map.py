"""Parity of the device variable map (pfs_map, pfs_download_map,
pfs_map_lits, and the packed pfs_download_cnf) with the reference's host
write-back after the rounds.  Integer work: bit-exact.

* golden: simplify each fixture on the GPU, map, and compare with the
  state the reference's own Solver::map(true) / newBeginning() left;
* a second simplify on the mapped resident formula equals the oracle run
  on the mapped state (the device vorg, vstate and assignments carry over);
* full-size Tseitin output: mapped clauses are the map of the unmapped ones;
* threshold, identity and error paths.
"""
import dataclasses
import os

import numpy as np
import pytest

from oracle import oracle, vmap
from parafrost_amd import gen
from parafrost_amd.simplifier import PfsError, Simplifier
from tests.common import GOLDEN, diff, load_manifest, oracle_kw, read_golden

pytestmark = pytest.mark.gpu
MAPPED = [e for e in load_manifest() if "map" in e]


def arena(res):
    """Result clauses as the reference's arena lists: originals, then learnts
    (flags: bit 0 learnt, bits 4-5 the learnt clause's usage)."""
    off = np.zeros(res.sizes.shape[0] + 1, np.int64)
    np.cumsum(res.sizes, out=off[1:])
    learnt = (res.flags & 1) != 0
    order = np.concatenate([np.nonzero(~learnt)[0], np.nonzero(learnt)[0]])
    lits = [res.lits[off[i]:off[i + 1]] for i in order]
    return (res.sizes[order], np.where(learnt[order], 1 | (res.flags[order] & 0x30), 0).astype(np.uint8),
            np.where(learnt[order], res.lbd[order], 0),
            np.concatenate(lits) if lits else np.zeros(0, np.uint32))


def simplified(entry):
    inp, _ = read_golden(entry)
    s = Simplifier(0, **entry["options"])
    s.upload_state(inp)
    st = s.simplify()
    return s, st, inp


@pytest.mark.parametrize("entry", MAPPED, ids=[e["name"] for e in MAPPED])
def test_map_matches_reference(entry):
    print("map golden", entry["name"], flush=True)
    ref = vmap.read_pfm(os.path.join(GOLDEN, entry["map"]))
    s, st, inp = simplified(entry)
    try:
        assert st == 2
        old = s.maxvar
        info = s.map()
        assert info["applied"] == ref["applied"]
        assert info["old_maxvar"] == old
        res = s.result(st)
        assert res.maxvar == ref["maxvar"]
        sizes, flags, lbd, lits = arena(res)
        assert np.array_equal(sizes, ref["sizes"])
        assert np.array_equal(flags, ref["flags"])
        assert np.array_equal(lbd, ref["lbd"])
        assert np.array_equal(lits, ref["lits"])
        assert np.array_equal(res.vstate, ref["vstate"])
        m = s.download_map(old)
        model = s.map_lits(2 * np.arange(old + 1, dtype=np.uint32))
        assert np.array_equal(model, ref["model_lits"])
        if ref["applied"]:
            assert info["new_maxvar"] == ref["maxvar"]
            assert ([info["trail_lit"]] if info["trail_lit"] else []) == ref["trail"].tolist()
            vorg = np.zeros(info["new_maxvar"] + 1, np.uint32)
            src = np.nonzero(m)[0]
            vorg[m[src]] = src
            assert np.array_equal(vorg, ref["vorg"])
        else:
            assert np.array_equal(m, np.arange(old + 1))
    finally:
        s.close()


@pytest.mark.parametrize("entry", [e for e in MAPPED if e["name"].endswith(".p1bce")],
                         ids=[e["name"] for e in MAPPED if e["name"].endswith(".p1bce")])
def test_simplify_after_map_matches_oracle(entry):
    """The resident formula after pfs_map simplified again equals the oracle
    on the mapped state (clauses, units, stack in original numbering)."""
    print("map then simplify", entry["name"], flush=True)
    ref = vmap.read_pfm(os.path.join(GOLDEN, entry["map"]))
    s, st, _ = simplified(entry)
    try:
        s.map(force=True)
        mid = s.result(st)  # units and stack words accumulate on a resident formula
        g = s.result(s.simplify())
    finally:
        s.close()
    g = dataclasses.replace(g, units=g.units[mid.units.shape[0]:], resolved=g.resolved[mid.resolved.shape[0]:])
    if not ref["applied"]:
        # the reference did not map this one: map it in the oracle too
        inp, out = read_golden(entry)
        trail = np.concatenate([inp.units, out.units]).astype(np.uint32)
        assign = np.zeros(out.maxvar + 1, np.uint32)
        assign[trail >> 1] = trail + 1
        ref = vmap.map_state(out.maxvar, out.vstate, assign, out.sizes, out.flags, out.lbd, out.lits, trail=trail,
                             force=True)
    off = np.zeros(ref["sizes"].shape[0] + 1, np.uint64)
    np.cumsum(ref["sizes"], out=off[1:])
    r = oracle.simplify(ref["maxvar"], off, ref["lits"], ref["flags"], ref["lbd"], ref["vstate"], None, ref["vorg"],
                        trail=ref["trail"], det=1, **oracle_kw(entry["options"]))
    d = diff(g, r)
    assert not d, d


def test_full_tseitin_map_round_trip():
    """Tseitin circuit, 200,000 gates: the mapped clauses are the map of the
    clauses downloaded before the map; the map is dense and ordered."""
    cnf, _ = gen.prepare(gen.tseitin(200_000, 500, 100, seed=0x5EED0002))
    s = Simplifier(0, phases=1, bce=1)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
        st = s.simplify()
        assert st == 2
        before = s.result(st)
        old = s.maxvar
        info = s.map(force=True)
        after = s.result(st)
        m = s.download_map(old)
    finally:
        s.close()
    kept = np.nonzero(m)[0]
    assert np.array_equal(m[kept], np.arange(1, kept.size + 1))
    assert info["new_maxvar"] == kept.size
    assert np.array_equal(after.sizes, before.sizes)
    assert np.array_equal(after.flags, before.flags)
    expect = (2 * m[before.lits >> 1] + (before.lits & 1)).astype(np.uint32)
    assert np.array_equal(after.lits, expect)
    # every clause literal belongs to a kept, unassigned variable
    assert np.all(m[before.lits >> 1] != 0)
    assert np.array_equal(after.vstate[m[kept]], before.vstate[kept])


def test_map_threshold_and_identity():
    cnf = gen.uniform_ksat(2000, 8520, 3, seed=5)
    s = Simplifier(0, mapperc=1.0)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        st = s.simplify()
        before = s.result(st)
        info = s.map()
        assert info["applied"] == 0 and info["new_maxvar"] == cnf.maxvar
        assert np.array_equal(s.download_map(cnf.maxvar), np.arange(cnf.maxvar + 1))
        after = s.result(st)
        assert np.array_equal(after.lits, before.lits)
        lits = np.array([2, 3, 0, 2 * cnf.maxvar + 1], np.uint32)
        assert np.array_equal(s.map_lits(lits), lits)
    finally:
        s.close()


def test_map_errors():
    cnf = gen.uniform_ksat(300, 1278, 3, seed=12)
    s = Simplifier(0, keep_input=True)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.map()
        st = s.simplify()
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.download_map(cnf.maxvar)
        s.map(force=True)
        with pytest.raises(PfsError, match="BAD_INPUT"):
            s.map_lits(np.array([2 * cnf.maxvar + 2], np.uint32))
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.rewind()
        assert st == 2
    finally:
        s.close()
    # an appended clause leaves root units pending: simplify before mapping
    s = Simplifier(0)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        s.simplify()
        s.append(cnf.maxvar + 1, np.array([0, 1], np.uint64), np.array([2 * cnf.maxvar + 2], np.uint32))
        with pytest.raises(PfsError, match="BAD_STATE"):
            s.map()
    finally:
        s.close()
