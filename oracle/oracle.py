"""ctypes wrapper of the oracle restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, asdict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OracleOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "phases", "mu_pos", "mu_neg", "lits_min", "ve_clause_limit", "xor_max_arity",
        "sub_limit", "bce_limit", "ere_limit", "ere_max_resolvent", "lcve_min", "shrink_rate", "lbd_tier1",
        "ve_en", "ve_plus_en", "sub_en", "bce_en", "ere_en", "all_en", "ve_lbound_en", "det", "proof_en",
        "lcve_max_occurs", "lcve_clause_max", "aggr_sort")]


# Defaults of the reference CPU simplifier options (src/cpu/options.cpp:24-46).
DEFAULTS = dict(phases=5, mu_pos=32, mu_neg=32, lits_min=500, ve_clause_limit=100, xor_max_arity=10,
                sub_limit=3000, bce_limit=3000, ere_limit=3000, ere_max_resolvent=1000, lcve_min=1,
                shrink_rate=2, lbd_tier1=2, ve_en=1, ve_plus_en=1, sub_en=1, bce_en=0, ere_en=1, all_en=0,
                ve_lbound_en=0, det=1, proof_en=0,
                # GPU-build election limits and the aggressive output sort: off
                lcve_max_occurs=2**31 - 1, lcve_clause_max=2**31 - 1, aggr_sort=0)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.oracle_run.restype = ctypes.c_void_p
        L.oracle_run.argtypes = [ctypes.POINTER(OracleOpts), ctypes.c_uint32, ctypes.c_uint64] + [ctypes.c_void_p] * 7 + [ctypes.c_uint32]
        L.oracle_run_trail.restype = ctypes.c_void_p
        L.oracle_run_trail.argtypes = ([ctypes.POINTER(OracleOpts), ctypes.c_uint32, ctypes.c_uint64]
                                       + [ctypes.c_void_p] * 8 + [ctypes.c_uint32, ctypes.c_uint32])
        L.oracle_status.argtypes = [ctypes.c_void_p]
        L.oracle_seconds.argtypes = [ctypes.c_void_p]
        L.oracle_seconds.restype = ctypes.c_double
        L.oracle_sizes.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_get.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 8
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_proof_size.argtypes = [ctypes.c_void_p]
        L.oracle_proof_size.restype = ctypes.c_uint64
        L.oracle_get_proof.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _LIB = L
    return _LIB


@dataclass
class Result:
    status: int
    maxvar: int
    sizes: np.ndarray
    flags: np.ndarray
    lbd: np.ndarray
    lits: np.ndarray
    units: np.ndarray
    resolved: np.ndarray
    vstate: np.ndarray
    elected: np.ndarray
    seconds: float
    proof: bytes = b""  # binary DRAT lines of the run (when proof_en)


def _p(a):
    return None if a is None else a.ctypes.data


def simplify(maxvar, offsets, lits, flags=None, lbd=None, vstate=None, assumed=None, vorg=None,
             nassigned=0, trail=None, npropagated=None, **opts) -> Result:
    """Runs the oracle.  Root assignments: either only their number
    (nassigned; the formula holds none of their variables) or the literals
    themselves (trail), of which the first npropagated are propagated and the
    rest still have to be.  The result's units are the new ones."""
    o = dict(DEFAULTS)
    o.update(opts)
    oo = OracleOpts(**o)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    lits = np.ascontiguousarray(lits, np.uint32)
    flags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    lbd = None if lbd is None else np.ascontiguousarray(lbd, np.uint32)
    vstate = None if vstate is None else np.ascontiguousarray(vstate, np.uint8)
    assumed = None if assumed is None else np.ascontiguousarray(assumed, np.uint8)
    vorg = None if vorg is None else np.ascontiguousarray(vorg, np.uint32)
    L = lib()
    if trail is None:
        h = L.oracle_run(ctypes.byref(oo), maxvar, offsets.shape[0] - 1, _p(offsets), _p(lits), _p(flags),
                         _p(lbd), _p(vstate), _p(assumed), _p(vorg), nassigned)
        skip = 0  # the oracle's trail starts empty: every unit is new
    else:
        trail = np.ascontiguousarray(trail, np.uint32)
        npr = trail.shape[0] if npropagated is None else int(npropagated)
        h = L.oracle_run_trail(ctypes.byref(oo), maxvar, offsets.shape[0] - 1, _p(offsets), _p(lits), _p(flags),
                               _p(lbd), _p(vstate), _p(assumed), _p(vorg), _p(trail), trail.shape[0], npr)
        skip = trail.shape[0]  # the given literals lead the oracle's trail
    try:
        sz = np.zeros(6, np.uint64)
        L.oracle_sizes(h, sz.ctypes.data)
        n, nl, nu, nr, nv, ne = (int(x) for x in sz)
        r = Result(L.oracle_status(h), maxvar, np.zeros(n, np.uint32), np.zeros(n, np.uint8),
                   np.zeros(n, np.uint32), np.zeros(nl, np.uint32), np.zeros(nu, np.uint32),
                   np.zeros(nr, np.uint32), np.zeros(nv, np.uint8), np.zeros(ne, np.uint32),
                   L.oracle_seconds(h))
        L.oracle_get(h, _p(r.sizes), _p(r.flags), _p(r.lbd), _p(r.lits), _p(r.units), _p(r.resolved),
                     _p(r.vstate), _p(r.elected))
        r.units = r.units[skip:]
        npf = int(L.oracle_proof_size(h))
        if npf:
            pf = np.zeros(npf, np.uint8)
            L.oracle_get_proof(h, pf.ctypes.data)
            r.proof = pf.tobytes()
        return r
    finally:
        L.oracle_free(h)


def simplify_state(st, **opts) -> Result:
    """Run on a PFD input state (parafrost_amd.pfd.State)."""
    off = np.zeros(st.sizes.shape[0] + 1, np.uint64)
    np.cumsum(st.sizes, out=off[1:])
    return simplify(st.maxvar, off, st.lits, st.flags, st.lbd, st.vstate, None, None,
                    nassigned=int(st.units.shape[0]), **opts)
