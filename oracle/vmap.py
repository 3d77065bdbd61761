"""Restatement of the reference's variable map (the host write-back after the
simplify rounds), in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, never
by the product package.

Follows src/cpu/vmap.h (VMAP) and src/cpu/vmap.cpp (Solver::map) of
ParaFROST, for the part that the simplifier's handoff touches: the map
itself, the clauses (newBeginning, src/cpu/simplify.cpp:251-269, with
VMAP::mapClause), the trail, vorg, the variable states and the model's
literal table.

Pinned by tests/test_map.py against the reference's own Solver::map(true)
run on every golden fixture (the PFM1 dumps written by
oracle/ref_harness.cpp), and by closed-form properties of the map.
"""
from __future__ import annotations

import numpy as np


def can_map(vstate: np.ndarray, maxvar: int, map_perc: float = 0.2) -> bool:
    """Solver::canMap (src/cpu/solver.h:164-169): more than map_perc of the
    variables inactive (maxInactive = melted + frozen + substituted,
    src/cpu/definitions.h:174)."""
    inactive = int(np.count_nonzero(vstate[1:maxvar + 1]))
    return inactive > map_perc * maxvar


def initiate(maxvar: int, vstate: np.ndarray, assign: np.ndarray):
    """VMAP::initiate (vmap.h:63-77).  assign[v] = true literal + 1, or 0.
    Returns (mapped, first, first_true): mapped[v] is the new variable of v
    (0: dropped); first is the first variable that is inactive and assigned
    (0: none) and first_true its new literal that is true."""
    v = np.arange(maxvar + 1, dtype=np.int64)
    cand = (v > 0) & (vstate[: maxvar + 1] != 0) & (assign[: maxvar + 1] != 0)
    first = int(np.argmax(cand)) if cand.any() else 0
    keep = (v > 0) & ((vstate[: maxvar + 1] == 0) | (v == first))
    mapped = np.where(keep, np.cumsum(keep), 0).astype(np.uint32)
    first_true = 0
    if first:
        true_lit = int(assign[first]) - 1
        # newLitDL0 = V2L(mappedFirstDL0), flipped when the positive literal is false
        first_true = 2 * int(mapped[first]) + (true_lit & 1)
    return mapped, first, first_true


def map_lits(lits: np.ndarray, mapped, first, first_true, assign) -> np.ndarray:
    """VMAP::mapLit (vmap.h:45-60) over an array; 0 entries stay 0
    (mapOrgs, vmap.h:112-131)."""
    lits = np.asarray(lits, np.int64)
    out = np.zeros(lits.shape, np.uint32)
    nz = lits != 0
    var = lits >> 1
    nv = np.where(nz, mapped[np.where(nz, var, 0)], 0).astype(np.int64)
    direct = nz & (nv != 0) & (var != first)
    out[direct] = (2 * nv[direct] + (lits[direct] & 1)).astype(np.uint32)
    rest = nz & ~direct
    a = assign[var[rest]].astype(np.int64)
    val = np.where(a == 0, 0, np.where(lits[rest] == a - 1, first_true, first_true ^ 1))
    out[rest] = val.astype(np.uint32)
    return out


def map_state(maxvar, vstate, assign, sizes, flags, lbd, lits, vorg=None, trail=None, model_lits=None,
              map_perc: float = 0.2, force: bool = False) -> dict:
    """The write-back after the rounds (simplify.cpp:227-228): Solver::map
    (vmap.cpp:61-135) when can_map holds (or force), else newBeginning
    without renumbering.  Clauses come in the simplifier's order and go out
    as the reference's arena lists: originals, then learnts
    (newClause, src/cpu/sclause.cpp:22-57).  Returns the fields of a PFM1
    dump."""
    vstate = np.asarray(vstate, np.uint8)
    assign = np.asarray(assign, np.uint32)
    sizes = np.asarray(sizes, np.int64)
    flags = np.asarray(flags, np.uint8)
    lbd = np.asarray(lbd, np.uint32)
    lits = np.asarray(lits, np.uint32)
    vorg = np.arange(maxvar + 1, dtype=np.uint32) if vorg is None else np.asarray(vorg, np.uint32)
    trail = np.zeros(0, np.uint32) if trail is None else np.asarray(trail, np.uint32)
    model_lits = (2 * np.arange(maxvar + 1, dtype=np.uint32) if model_lits is None
                  else np.asarray(model_lits, np.uint32))
    off = np.zeros(sizes.shape[0] + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    learnt = (flags & 1) != 0
    order = np.concatenate([np.nonzero(~learnt)[0], np.nonzero(learnt)[0]])
    applied = bool(force) or can_map(vstate, maxvar, map_perc)
    if applied:
        mapped, first, first_true = initiate(maxvar, vstate, assign)
        new_maxvar = int(mapped.max()) if mapped.size else 0
        lits = map_lits(lits, mapped, first, first_true, assign)
        model_lits = map_lits(model_lits, mapped, first, first_true, assign)
        # mapShrinkLits (vmap.h:158-169) on the trail
        tv = trail >> 1
        keep = mapped[tv] != 0
        trail = (2 * mapped[tv[keep]] + (trail[keep] & 1)).astype(np.uint32)
        # mapShrinkVars (vmap.h:145-156) on vorg; mapSP copies vstate (vmap.h:80-100)
        new_vorg = np.zeros(new_maxvar + 1, np.uint32)
        new_vs = np.zeros(new_maxvar + 1, np.uint8)
        src = np.nonzero(mapped)[0]
        new_vorg[mapped[src]] = vorg[src]
        new_vs[mapped[src]] = vstate[src]
        new_vorg[0] = vorg[0]
        vorg, vstate, maxvar = new_vorg, new_vs, new_maxvar
    else:
        vstate = vstate.copy()
        vstate[0] = 0
    out_sizes = sizes[order]
    out_lits = np.concatenate([lits[off[i]:off[i + 1]] for i in order]) if order.size else np.zeros(0, np.uint32)
    return dict(applied=int(applied), maxvar=int(maxvar), vorg=vorg, trail=trail, model_lits=model_lits,
                vstate=vstate[: maxvar + 1], sizes=out_sizes.astype(np.uint32),
                flags=np.where(learnt[order], 1 | (flags[order] & 0x30), 0).astype(np.uint8), lbd=np.where(learnt[order], lbd[order], 0).astype(np.uint32),
                lits=out_lits.astype(np.uint32))


def read_pfm(path_or_bytes) -> dict:
    """Reads a PFM1 dump (oracle/ref_harness.cpp write_map)."""
    import gzip
    import struct
    if isinstance(path_or_bytes, (bytes, bytearray)):
        b = bytes(path_or_bytes)
    else:
        op = gzip.open if str(path_or_bytes).endswith(".gz") else open
        with op(path_or_bytes, "rb") as f:
            b = f.read()
    if b[:4] != b"PFM1":
        raise ValueError("not a PFM1 dump")
    o = 4
    applied, maxvar = struct.unpack_from("<II", b, o)
    o += 8

    def arr(dt, n=None):
        nonlocal o
        if n is None:
            (n,) = struct.unpack_from("<Q", b, o)
            o += 8
        a = np.frombuffer(b, dt, n, o).copy()
        o += n * np.dtype(dt).itemsize
        return a

    vorg = arr("<u4")
    trail = arr("<u4")
    model_lits = arr("<u4")
    vstate = arr("u1")
    (nc,) = struct.unpack_from("<Q", b, o)
    o += 8
    sizes = arr("<u4", nc)
    flags = arr("u1", nc)
    lbd = arr("<u4", nc)
    lits = arr("<u4")
    return dict(applied=applied, maxvar=maxvar, vorg=vorg, trail=trail, model_lits=model_lits, vstate=vstate,
                sizes=sizes, flags=flags, lbd=lbd, lits=lits)
