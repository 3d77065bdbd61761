"""The ELS restatement (oracle/els.py) against the reference.

* the reference's own Solver::decompose rounds + rootify + shrinkTop
  (PFREF_ELS_OUT of oracle/ref_harness.cpp) on the fixtures: planted
  equivalence classes with chords, units, clauses that turn into units
  once substituted, a refuting class, one and five rounds, 2-SAT, and the
  generated uniform, Tseitin, 2-SAT and heavy-tailed texts;
* the SCC routine against reachability computed by brute force;
* the stack records extend every model of the result to a model of the
  input (truth tables on tiny formulas);
* planted mistakes in the restatement are caught by the fixtures.
"""
import itertools
import json
import os
import types

import numpy as np
import pytest

from oracle import els
from parafrost_amd import gen, pfd
from tests.common import GOLDEN, extend_model

with open(os.path.join(GOLDEN, "els_manifest.json")) as f:
    EMANIFEST = json.load(f)


def clause_lists(st):
    off = np.zeros(st.sizes.shape[0] + 1, np.int64)
    np.cumsum(st.sizes, out=off[1:])
    return [st.lits[off[k]:off[k + 1]].tolist() for k in range(st.sizes.shape[0])]


def run_oracle(mod, entry):
    inp = pfd.read(os.path.join(GOLDEN, entry["input"]))
    return mod.els(inp.maxvar, clause_lists(inp), [bool(f & 1) for f in inp.flags], inp.units.tolist(),
                   inp.vstate.tolist(), rounds=entry["rounds"] or 2)


def agrees(mod, entry):
    ref = pfd.read(os.path.join(GOLDEN, entry["output"]))
    st, s = run_oracle(mod, entry)
    if st != ref.status:
        return False
    if not st:
        return True
    got = [sorted(c) for c in s.clauses if c is not None]
    return (got == [sorted(c) for c in clause_lists(ref)] and s.stack == ref.resolved.tolist()
            and sorted(s.trail) == sorted(ref.units.tolist()) and s.vstate == ref.vstate.tolist())


def test_fixtures_cover_the_branches():
    refs = [pfd.read(os.path.join(GOLDEN, e["output"])) for e in EMANIFEST]
    assert {r.status for r in refs} == {0, 2}
    assert sum(int((r.vstate == 4).sum()) for r in refs) > 5000
    # substitution units: the trail grows during ELS in some fixture
    grew = [pfd.read(os.path.join(GOLDEN, e["input"])).units.shape[0] < r.units.shape[0]
            for e, r in zip(EMANIFEST, refs) if r.status]
    assert any(grew)


@pytest.mark.parametrize("entry", EMANIFEST, ids=[e["name"] for e in EMANIFEST])
def test_oracle_matches_reference(entry):
    assert agrees(els, entry)


@pytest.mark.parametrize("seed", range(20))
def test_scc_against_reachability(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 24))
    adj = [[] for _ in range(n)]
    for _ in range(int(rng.integers(n, 3 * n))):
        a, b = (int(x) for x in rng.integers(2, n, 2))
        adj[a].append(b)
    act = {x for x in range(2, n) if rng.random() < 0.9}
    comp, comps = els._sccs(n, adj, lambda l: l in act)
    reach = np.eye(n, dtype=bool)
    for a in range(n):
        for b in adj[a]:
            if a in act and b in act:
                reach[a, b] = True
    for k in range(n):
        reach |= reach[:, k:k + 1] & reach[k:k + 1, :]
    for a in act:
        for b in act:
            assert (comp[a] == comp[b]) == bool(reach[a, b] and reach[b, a])
    assert all(comp[x] == -1 for x in range(n) if x not in act)


@pytest.mark.parametrize("seed", range(12))
def test_stack_extends_models(seed):
    """Every model of the ELS result, extended over the stack records, is a
    model of the input (and the result is satisfiable iff the input is)."""
    cnf = gen.equivalences(10, 6, 2, 3, seed=seed, anti=int(seed % 3 == 0))
    maxvar = cnf.maxvar
    cls = [cnf.lits[int(cnf.offsets[i]):int(cnf.offsets[i + 1])].tolist() for i in range(cnf.nclauses)]
    units = [c[0] for c in cls if len(c) == 1]
    body = [c for c in cls if len(c) > 1]
    vstate = [0] * (maxvar + 1)
    for u in units:
        vstate[u >> 1] = 2
    if len({u >> 1 for u in units}) < len(units):
        return
    st, s = els.els(maxvar, body, [False] * len(body), units, vstate)

    def sat(clauses, vals):
        return all(any(vals[x >> 1] == (x & 1 == 0) for x in c) for c in clauses)

    inp_sat = any(sat(cls, [None] + list(bits)) for bits in itertools.product([False, True], repeat=maxvar))
    if st == 0:
        assert not inp_sat
        return
    out = [c for c in s.clauses if c is not None] + [[u] for u in s.trail]
    found = False
    for bits in itertools.product([False, True], repeat=maxvar):
        vals = [None] + list(bits)
        if not sat(out, vals):
            continue
        found = True
        ext = extend_model({v: vals[v] for v in range(1, maxvar + 1)}, np.array(s.stack, np.uint32))
        assert sat(cls, [None] + [ext[v] for v in range(1, maxvar + 1)])
    assert found == inp_sat


MUTANTS = {
    "largest representative": ("            rep = min(members, key=lambda l: l >> 1)\n",
                               "            rep = max(members, key=lambda l: l >> 1)\n"),
    "units seen only after the pass": ("                    self.enqueue(out[0])\n",
                                       "                    self.trail.append(out[0])\n"),
    "witness sign": ("self.org(p), self.org(other ^ 1), 2, self.org(p ^ 1), 1]",
                     "self.org(p), self.org(other), 2, self.org(p ^ 1), 1]"),
    "no propagation after the pass": ("        if not self.propagate():\n", "        if False:\n"),
    "substituted although the representative is fixed": ("                if not self.vstate[other >> 1]:\n",
                                                          "                if True:\n"),
    "tautologies kept": ("                            elif m != (other & 1):\n                                satisfied = True\n",
                         "                            elif m != (other & 1):\n                                pass\n"),
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(mutant):
    src = open(els.__file__).read()
    old, new = MUTANTS[mutant]
    assert old in src
    mod = types.ModuleType("els_mutant")
    exec(compile(src.replace(old, new, 1), "els_mutant", "exec"), mod.__dict__)
    assert any(not agrees(mod, e) for e in EMANIFEST)
