"""A small forward DRAT checker for the tests (RUP lines only).

Every added line must be a reverse-unit-propagation consequence of the
formula so far: assigning the negation of its literals and propagating
units must reach a conflict.  The simplifier's lines are RUP by
construction (resolvents, self-subsuming strengthening, substitution through
the gate's binaries, units), and it deletes only clauses it proves
redundant (subsumed, equal to a resolvent), so no RAT check is needed.
Deletions of clauses that are the reason of a top-level unit are applied to
the clause set only, as DRAT-trim does for unit deletions.

Literals are the binary DRAT numbers 2*v + sign, the simplifier's own
encoding when variables are not renumbered.
"""
from __future__ import annotations


def parse(proof: bytes):
    """Binary DRAT bytes -> [(kind, (lit, ...))], kind 'a' or 'd'."""
    out = []
    i, n = 0, len(proof)
    while i < n:
        kind = chr(proof[i])
        if kind not in "ad":
            raise ValueError(f"bad line tag {proof[i]!r} at byte {i}")
        i += 1
        lits = []
        while True:
            r = sh = 0
            while True:
                if i >= n:
                    raise ValueError("truncated proof")
                x = proof[i]
                i += 1
                r |= (x & 127) << sh
                sh += 7
                if x < 128:
                    break
            if r == 0:
                break
            lits.append(r)
        out.append((kind, tuple(lits)))
    return out


class Checker:
    """Clause set with two-watched-literal unit propagation."""

    def __init__(self, nvars: int):
        self.val = [0] * (2 * nvars + 2)  # per literal: 1 true, -1 false, 0 unassigned
        self.trail: list[int] = []
        self.clauses: list[list[int]] = []
        self.alive: list[bool] = []
        self.watch: dict[int, list[int]] = {}
        self.index: dict[tuple, list[int]] = {}
        self.qhead = 0
        self.inconsistent = False

    def _assign(self, lit):
        self.val[lit] = 1
        self.val[lit ^ 1] = -1
        self.trail.append(lit)

    def _propagate(self) -> bool:
        """Returns False on a conflict."""
        while self.qhead < len(self.trail):
            lit = self.trail[self.qhead]
            self.qhead += 1
            false = lit ^ 1
            ws = self.watch.get(false, [])
            keep = []
            conflict = False
            i = 0
            while i < len(ws):
                ci = ws[i]
                i += 1
                if not self.alive[ci]:
                    continue
                c = self.clauses[ci]
                if c[0] == false:
                    c[0], c[1] = c[1], c[0]
                if self.val[c[0]] == 1:
                    keep.append(ci)
                    continue
                for k in range(2, len(c)):
                    if self.val[c[k]] != -1:
                        c[1], c[k] = c[k], c[1]
                        self.watch.setdefault(c[1], []).append(ci)
                        break
                else:
                    keep.append(ci)
                    if self.val[c[0]] == -1:
                        conflict = True
                        keep.extend(ws[i:])
                        break
                    self._assign(c[0])
            self.watch[false] = keep
            if conflict:
                return False
        return True

    def add(self, lits) -> None:
        lits = list(dict.fromkeys(lits))
        key = tuple(sorted(lits))
        ci = len(self.clauses)
        self.clauses.append(lits)
        self.alive.append(True)
        self.index.setdefault(key, []).append(ci)
        if self.inconsistent:
            return
        if not lits:
            self.inconsistent = True
            return
        if len(lits) == 1:
            if self.val[lits[0]] == -1:
                self.inconsistent = True
            elif self.val[lits[0]] == 0:
                self._assign(lits[0])
                if not self._propagate():
                    self.inconsistent = True
            return
        # watch two literals that are not false if possible
        lits.sort(key=lambda l: self.val[l] == -1)
        self.watch.setdefault(lits[0], []).append(ci)
        self.watch.setdefault(lits[1], []).append(ci)
        if self.val[lits[0]] == -1:
            self.inconsistent = True
        elif self.val[lits[1]] == -1 and self.val[lits[0]] == 0:
            self._assign(lits[0])
            if not self._propagate():
                self.inconsistent = True

    def delete(self, lits) -> bool:
        key = tuple(sorted(dict.fromkeys(lits)))
        ids = self.index.get(key)
        if not ids:
            return False
        self.alive[ids.pop()] = False
        return True

    def rup(self, lits) -> bool:
        if self.inconsistent:
            return True
        mark = len(self.trail)
        head = self.qhead
        ok = False
        for l in lits:
            if self.val[l] == 1:
                ok = True
                break
            if self.val[l] == 0:
                self._assign(l ^ 1)
        if not ok:
            ok = not self._propagate()
        for l in self.trail[mark:]:
            self.val[l] = self.val[l ^ 1] = 0
        del self.trail[mark:]
        self.qhead = head
        return ok


def check(nvars: int, clauses, units, proof: bytes, expect_empty: bool = False) -> tuple[bool, str]:
    """Checks 'proof' against the formula 'clauses' (iterable of literal
    sequences) plus the root 'units'.  Returns (ok, message)."""
    ch = Checker(nvars)
    for u in units:
        ch.add([int(u)])
    for c in clauses:
        ch.add([int(x) for x in c])
    has_empty = False
    for n, (kind, lits) in enumerate(parse(proof)):
        if kind == "d":
            if not ch.delete(lits):
                return False, f"line {n}: deleted clause {lits} is not in the formula"
            continue
        if not ch.rup(lits):
            return False, f"line {n}: clause {lits} is not RUP"
        if not lits:
            has_empty = True
        ch.add(lits)
    if expect_empty and not has_empty:
        return False, "no empty clause"
    return True, "ok"
