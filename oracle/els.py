"""Restatement of the reference's equivalent-literal substitution (ELS:
SCC decomposition of the binary implication graph and substitution), in
pure Python (small inputs).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows ParaFROST src/cpu/equivalence.cpp: Solver::decompose (:41-199:
Tarjan over the binary watches, each SCC represented by its literal of
smallest variable, assumed variables kept, a literal and its negation in
one SCC refute), Solver::substitute (:201-306: learnts, then originals;
per clause the literals mapped to their representatives, duplicates
dropped, tautologies and root-satisfied clauses removed, false literals
dropped, units enqueued at once and seen by the clauses after them), the
propagation after it, markSubstituted (solver.h:542-549) and
MODEL::saveBinary (model.cpp:271-288) for the eliminated-variable stack;
ELS (:316-323) repeats decompose while it succeeds.  The scenario is the
one oracle/ref_harness.cpp runs (PFREF_ELS_OUT): a root-clean formula,
decompose rounds, then rootify + shrinkTop.

Pinned by tests/test_els.py against those reference runs on formulas with
planted equivalence classes (with chords, units, a refuting class) and on
the golden texts, by closed-form checks of the SCC representatives, and
by planted mistakes.
"""
from __future__ import annotations


def _sccs(nlits, adj, active):
    """Strongly connected components of the literal graph restricted to
    active literals (iterative Tarjan).  Returns comp[l] (component id) for
    active literals, -1 otherwise, and the list of components."""
    index = [0] * nlits
    low = [0] * nlits
    on = [False] * nlits
    comp = [-1] * nlits
    comps = []
    stack = []
    counter = 1
    for root in range(2, nlits):
        if not active(root) or index[root]:
            continue
        work = [(root, 0)]
        index[root] = low[root] = counter
        counter += 1
        stack.append(root)
        on[root] = True
        while work:
            v, i = work[-1]
            succ = adj[v]
            if i < len(succ):
                work[-1] = (v, i + 1)
                w = succ[i]
                if not active(w):
                    continue
                if not index[w]:
                    index[w] = low[w] = counter
                    counter += 1
                    stack.append(w)
                    on[w] = True
                    work.append((w, 0))
                elif on[w]:
                    low[v] = min(low[v], index[w])
            else:
                work.pop()
                if work:
                    u = work[-1][0]
                    low[u] = min(low[u], low[v])
                if low[v] == index[v]:
                    members = []
                    while True:
                        w = stack.pop()
                        on[w] = False
                        comp[w] = len(comps)
                        members.append(w)
                        if w == v:
                            break
                    comps.append(members)
    return comp, comps


class State:
    """clauses: list of [lits] (None when deleted); learnt: list of bool;
    value[v]: true literal or None; vstate[v]: 0 active, 1 melted, 2 frozen,
    4 substituted; assumed: set of variables; vorg[v]."""

    def __init__(self, maxvar, clauses, learnt, trail, vstate, assumed=(), vorg=None):
        self.maxvar = maxvar
        self.clauses = [list(c) for c in clauses]
        self.learnt = list(learnt)
        self.value = [None] * (maxvar + 1)
        self.trail = list(trail)
        for u in trail:
            self.value[u >> 1] = u
        self.vstate = list(vstate)
        self.assumed = set(assumed)
        self.vorg = list(vorg) if vorg is not None else list(range(maxvar + 1))
        self.stack = []
        self.refuted = False

    def val(self, lit):
        v = self.value[lit >> 1]
        return None if v is None else v == lit

    def enqueue(self, lit):
        self.value[lit >> 1] = lit
        self.vstate[lit >> 1] = 2
        self.trail.append(lit)

    def org(self, lit):
        return 2 * self.vorg[lit >> 1] + (lit & 1)

    # ---- Solver::decompose
    def decompose(self) -> bool:
        n = 2 * (self.maxvar + 1)
        adj = [[] for _ in range(n)]
        for c in self.clauses:
            if c is not None and len(c) == 2:
                a, b = c
                adj[a ^ 1].append(b)
                adj[b ^ 1].append(a)

        def active(l):
            return self.vstate[l >> 1] == 0

        comp, comps = _sccs(n, adj, active)
        smallest = [0] * n  # 0 for inactive literals, as the reference's fresh array
        substituted = 0
        for members in comps:
            vars_ = [l >> 1 for l in members]
            if len(set(vars_)) < len(vars_):  # l and its negation in one SCC
                self.refuted = True
                return True
            rep = min(members, key=lambda l: l >> 1)
            for l in members:
                if (l >> 1) in self.assumed:
                    smallest[l] = l
                else:
                    smallest[l] = rep
                    substituted += l != rep
        units = binaries = False
        if substituted:
            order = [i for i, c in enumerate(self.clauses) if c is not None and self.learnt[i]] + \
                    [i for i, c in enumerate(self.clauses) if c is not None and not self.learnt[i]]
            reduced = []
            for i in order:
                c = self.clauses[i]
                if all(smallest[l] == l for l in c):
                    continue
                out, marks, satisfied = [], {}, False
                for lit in c:
                    x = self.val(lit)
                    if x is None:
                        other = smallest[lit]
                        y = self.val(other)
                        if y is None:
                            m = marks.get(other >> 1)
                            if m is None:
                                marks[other >> 1] = other & 1
                                out.append(other)
                            elif m != (other & 1):
                                satisfied = True
                        elif y:
                            satisfied = True
                    elif x:
                        satisfied = True
                    if satisfied:
                        break
                if satisfied:
                    reduced.append(i)
                elif not out:
                    self.refuted = True
                    return True
                elif len(out) == 1:
                    self.enqueue(out[0])
                    self.clauses[i] = None
                    units = True
                else:
                    binaries |= len(out) == 2
                    self.clauses[i] = out
            for i in reduced:
                self.clauses[i] = None
        if not self.propagate():
            self.refuted = True
            return True
        for v in range(1, self.maxvar + 1):
            if self.vstate[v]:
                continue
            p = 2 * v
            other = smallest[p]
            if other != p:
                if not self.vstate[other >> 1]:
                    self.vstate[v] = 4
                # MODEL::saveBinary(p, -other)
                self.stack += [self.org(p), self.org(other ^ 1), 2, self.org(p ^ 1), 1]
        return bool(substituted) and (units or binaries)

    def propagate(self) -> bool:
        """Root unit propagation to the fixpoint (assignments only)."""
        changed = True
        while changed:
            changed = False
            for c in self.clauses:
                if c is None or any(self.val(x) for x in c):
                    continue
                free = [x for x in c if self.val(x) is None]
                if not free:
                    return False
                if len(free) == 1:
                    self.enqueue(free[0])
                    changed = True
        return True

    def shrink(self):
        """rootify + shrinkTop: satisfied clauses removed, false literals dropped."""
        for i, c in enumerate(self.clauses):
            if c is None:
                continue
            if any(self.val(x) for x in c):
                self.clauses[i] = None
            else:
                self.clauses[i] = [x for x in c if self.val(x) is None]


def els(maxvar, clauses, learnt, trail, vstate, rounds=2, assumed=(), vorg=None):
    """ELS rounds then the top-level shrink.  Returns (status, State):
    status 0 when refuted, 2 otherwise."""
    s = State(maxvar, clauses, learnt, trail, vstate, assumed, vorg)
    for _ in range(rounds):
        ok = s.decompose()
        if s.refuted or not ok:
            break
    if s.refuted:
        return 0, s
    s.shrink()
    return 2, s
