"""Restatement of the reference's DIMACS parser and the root-level cleanup
that precedes the first simplify, in pure Python (small inputs).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows ParaFROST src/cpu/dimacs.cpp (Solver::parse, the file path
:81-110, and toClause :113-170) with the lexer of src/cpu/dimacs.h
(isSpace :33, eatLine :99, toInteger :101-111), then the constructor's root
BCP (src/cpu/solver.cpp:49) and rootify/shrinkTop (src/cpu/solver.h:196-201,
src/cpu/shrink.cpp:152-216): satisfied clauses removed, false literals
dropped.  The trail comes out as parse-time units in text order followed
by propagated units in breadth-first order (the reference's watch order
is not restated, so only the trail's set is comparable).

Pinned by tests/test_dimacs.py against the reference's own parse + BCP +
shrinkTop on the golden fixtures (<case>.in.pfd.gz, written by
oracle/ref_harness.cpp from the same texts) and on hand-made texts with
comments, duplicate literals, tautologies, units, '+' signs, blank
variants and '%' terminators, plus known-answer and error cases.
"""
from __future__ import annotations


class ParseError(ValueError):
    pass


def _blank(c: int) -> bool:
    return 9 <= c <= 13 or c == 32


def _digit(c: int) -> bool:
    return 48 <= c <= 57


def parse(text: bytes):
    """Returns (status, maxvar, header_clauses, clauses, trail): status 0
    when the text refutes itself (empty clause, opposite units, conflict in
    root propagation), else 2.  clauses: lists of literals 2v+sign after the
    cleanup, in text order."""
    n = len(text)
    i = 0
    maxvar, ncls = 0, 0
    header = False
    value: list = []
    orgs: list = []
    trail: list = []
    refuted = False

    def to_int():  # dimacs.h:101-111
        nonlocal i
        while i < n and _blank(text[i]):
            i += 1
        sign = 0
        if i < n and text[i] == ord("-"):
            sign, i = 1, i + 1
        elif i < n and text[i] == ord("+"):
            i += 1
        if i >= n or not _digit(text[i]):
            raise ParseError("expected a digit")
        v = 0
        while i < n and _digit(text[i]):
            v = (v * 10 + text[i] - 48) & 0xFFFFFFFF
            i += 1
        return v, sign

    def val(lit):
        v = value[lit >> 1]
        return None if v is None else (v == lit)

    while i < n and not refuted:
        while i < n and _blank(text[i]):
            i += 1
        if i >= n or text[i] == 0 or text[i] == ord("%"):
            break
        c = text[i]
        if c == ord("c"):  # eatLine
            while i < n and text[i] != 0:
                i += 1
                if text[i - 1] == ord("\n"):
                    break
            continue
        if c == ord("p"):
            if header:
                raise ParseError("second header")
            if text[i:i + 5] != b"p cnf":
                raise ParseError("header has wrong format")
            i += 5
            maxvar, s = to_int()
            if s:
                raise ParseError("number of variables in header is negative")
            if maxvar == 0:
                raise ParseError("zero number of variables in header")
            if maxvar >= 0x7FFFFFFE:
                raise ParseError("number of variables not supported")
            ncls, s = to_int()
            if s:
                raise ParseError("number of clauses in header is negative")
            if ncls == 0:
                raise ParseError("zero number of clauses in header")
            header = True
            value = [None] * (maxvar + 1)
            continue
        # toClause (dimacs.cpp:113-170)
        cl, marks, satisfied = [], {}, False
        while True:
            v, s = to_int()
            if v == 0:
                break
            if v > maxvar:
                raise ParseError("too many variables")
            lit = 2 * v + s
            m = marks.get(v)
            if m is None:
                marks[v] = s
                x = val(lit)
                if x is None:
                    cl.append(lit)
                elif x:
                    satisfied = True
            elif m != s:
                satisfied = True  # tautology
        if satisfied:
            continue
        if not cl:
            refuted = True
        elif len(cl) == 1:
            value[cl[0] >> 1] = cl[0]  # enqueueUnit
            trail.append(cl[0])
        elif len(orgs) + 1 > ncls:
            raise ParseError("too many clauses")
        else:
            orgs.append(cl)
    if not header:
        raise ParseError("no header")
    if refuted:
        return 0, maxvar, ncls, [], trail
    # root propagation (BCP to the fixpoint), then shrinkTop
    occurs: dict = {}
    for k, cl in enumerate(orgs):
        for lit in cl:
            occurs.setdefault(lit, []).append(k)
    head = 0
    while head < len(trail):
        u = trail[head]
        head += 1
        for k in occurs.get(u ^ 1, ()):
            cl = orgs[k]
            if any(val(x) for x in cl):
                continue
            free = [x for x in cl if val(x) is None]
            if not free:
                return 0, maxvar, ncls, [], trail
            if len(free) == 1:
                value[free[0] >> 1] = free[0]
                trail.append(free[0])
    out = []
    for cl in orgs:
        if any(val(x) for x in cl):
            continue
        out.append([x for x in cl if val(x) is None])
    return 2, maxvar, ncls, out, trail
