"""Helpers shared by the tests: golden fixtures, option names, comparisons."""
import json
import os

import numpy as np

from parafrost_amd import pfd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# reference option name -> oracle keyword (oracle/oracle.py OracleOpts)
ORACLE_KW = {"phases": "phases", "mupos": "mu_pos", "muneg": "mu_neg", "bce": "bce_en", "all": "all_en",
             "ve": "ve_en", "ve+": "ve_plus_en", "sub": "sub_en", "ere": "ere_en", "velbound": "ve_lbound_en",
             "veclausemax": "ve_clause_limit", "xormaxarity": "xor_max_arity", "submax": "sub_limit",
             "bcemax": "bce_limit", "eremax": "ere_limit", "ereclausemax": "ere_max_resolvent",
             "lcvemin": "lcve_min", "gcfreq": "shrink_rate", "vephasemin": "lits_min",
             "electionsmax": "lcve_max_occurs", "lcveclausemax": "lcve_clause_max", "aggresivesort": "aggr_sort"}


def oracle_kw(opts):
    kw = {ORACLE_KW[k]: int(v) for k, v in opts.items()}
    # ve is "ve || ve+" in the reference (src/cpu/options.cpp:282)
    if "ve_en" in kw or "ve_plus_en" in kw:
        kw["ve_en"] = int(bool(kw.get("ve_en", 1)) or bool(kw.get("ve_plus_en", 1)))
    return kw


def load_manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def read_golden(entry):
    return pfd.read(os.path.join(GOLDEN, entry["input"])), pfd.read(os.path.join(GOLDEN, entry["output"]))


def diff(a, b, units_as_set=True):
    """Differences between two simplifier results (objects with status, sizes,
    flags, lbd, lits, units, resolved, vstate)."""
    d = []
    if a.status != b.status:
        return [f"status {a.status} != {b.status}"]
    if a.status == 0:
        return d
    for k in ("sizes", "flags", "lbd", "lits", "resolved", "vstate"):
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        if x.shape != y.shape:
            d.append(f"{k}: shape {x.shape} != {y.shape}")
        elif not np.array_equal(x, y):
            i = int(np.nonzero(x != y)[0][0])
            d.append(f"{k}: first difference at {i}: {x[i]} != {y[i]}")
    ua, ub = np.asarray(a.units), np.asarray(b.units)
    if units_as_set:
        ua, ub = np.sort(ua), np.sort(ub)
    if ua.shape != ub.shape or not np.array_equal(ua, ub):
        d.append(f"units differ ({ua.shape[0]} vs {ub.shape[0]})")
    return d


def extend_model(values, resolved):
    """Model reconstruction over the eliminated-variable stack: records are
    [lits (witness first)..., size], replayed from the end; a record with no
    satisfied literal sets its witness true (src/cpu/model.cpp:101-162).
    values: dict var -> bool, updated in place."""
    i = len(resolved) - 1
    while i > 0:
        k = int(resolved[i])
        lits = [int(x) for x in resolved[i - k:i]]
        i -= k + 1
        sat = any(values.get(l >> 1, False) == (not (l & 1)) for l in lits)
        if not sat:
            w = lits[0]
            values[w >> 1] = not (w & 1)
    return values


def read_golden_proof(entry):
    import gzip
    with gzip.open(os.path.join(GOLDEN, entry["proof"]), "rb") as f:
        return f.read()


def clause_list(st):
    """Clauses of a PFD input state as lists of literals."""
    off = [0] + np.cumsum(st.sizes.astype(np.int64)).tolist()
    return [st.lits[off[i]:off[i + 1]].tolist() for i in range(st.sizes.shape[0])]


def proof_lines_canonical(proof: bytes):
    """Proof lines with each run of consecutive unit lines sorted: unit
    propagation finds the same units in any order, so its lines are compared
    as a set (the parallel propagation claims units level by level)."""
    from tests.drat import parse
    out, run = [], []
    for line in parse(proof):
        if line[0] == "a" and len(line[1]) == 1:
            run.append(line)
            continue
        out.extend(sorted(run))
        run = []
        out.append(line)
    out.extend(sorted(run))
    return out


def append_clean(assign, offsets, lits):
    """The clause clean-up of pfs_append (and of the reference's incremental
    itoClause, src/cpu/addinc.cpp:85-160) on the host: literals false under
    'assign' (dict variable -> true literal) are dropped, satisfied clauses
    and tautologies skipped, unit clauses assigned in order.  Returns (kept
    clauses, new unit literals, refuted); 'assign' is updated."""
    kept, units, refuted = [], [], False
    off = offsets.tolist()
    for i in range(len(off) - 1):
        cl, sat = [], False
        for x in lits[off[i]:off[i + 1]].tolist():
            a = assign.get(x >> 1)
            if a is not None:
                sat |= a == x
                continue
            if x ^ 1 in cl:
                sat = True
            if x not in cl:
                cl.append(x)
        if sat:
            continue
        if not cl:
            refuted = True
        elif len(cl) == 1:
            if (cl[0] >> 1) not in assign:
                assign[cl[0] >> 1] = cl[0]
                units.append(cl[0])
        else:
            kept.append(cl)
    return kept, units, refuted


def proof_digest(proof: bytes) -> str:
    """sha256 of a binary DRAT proof in canonical form: each run of
    consecutive unit-addition lines sorted (the parallel propagation claims
    units level by level, so their order within a run is free), and, in a
    proof that ends with the empty clause, the run of unit lines right
    before it left out: a propagation that ends in a conflict writes the
    units it assigned before meeting it, which depends on the order it
    visits clauses in (the sequential propagation writes them one by one,
    the level-parallel one stops at the level); they are implied by unit
    propagation and carry nothing the empty clause's check needs.  A line
    ends at its 0 byte: every literal's last varint byte is nonzero."""
    import hashlib
    h = hashlib.sha256()
    run = []
    lines = proof.split(b"\0")[:-1]
    if lines and lines[-1] == b"a":  # the empty clause: drop the unit run before it
        k = len(lines) - 1
        while k > 0 and lines[k - 1][:1] == b"a" and len(lines[k - 1]) > 1 and \
                all(b >= 128 for b in lines[k - 1][1:-1]):
            k -= 1
        lines = lines[:k] + lines[-1:]
    for line in lines:
        if line[:1] == b"a" and len(line) > 1 and all(b >= 128 for b in line[1:-1]):  # one literal
            run.append(line)
            continue
        for u in sorted(run):
            h.update(u + b"\0")
        run = []
        h.update(line + b"\0")
    for u in sorted(run):
        h.update(u + b"\0")
    return h.hexdigest()


def result_digest(r) -> dict:
    """sha256 digests of a simplifier result's arrays (units as a sorted
    set, proof in canonical form) and its sizes: the form in which
    full-size outputs are kept as fixtures."""
    import hashlib

    def h(a, dt):
        return hashlib.sha256(np.ascontiguousarray(np.asarray(a, dt)).tobytes()).hexdigest()

    out = {"status": int(r.status)}
    if r.status != 0:
        out.update({"clauses": int(np.asarray(r.sizes).shape[0]), "literals": int(np.asarray(r.lits).shape[0]),
                    "stack_words": int(np.asarray(r.resolved).shape[0]), "units": int(np.asarray(r.units).shape[0]),
                    "sizes": h(r.sizes, np.uint32), "flags": h(r.flags, np.uint8), "lbd": h(r.lbd, np.uint32),
                    "lits": h(r.lits, np.uint32), "resolved": h(r.resolved, np.uint32),
                    "vstate": h(r.vstate, np.uint8), "units_set": h(np.sort(np.asarray(r.units, np.uint32)), np.uint32)})
    if getattr(r, "proof", b""):
        out["proof_bytes"] = len(r.proof)
        out["proof"] = proof_digest(r.proof)
    return out
