"""ctypes wrapper of the test-only CDCL solver and model checker
(tests/cdcl/libcdcl.so, built from cdcl.cpp by its Makefile).

check_equisat() is the check run on simplifier outputs: the output formula
(clauses plus every root unit) is solved; a model is extended over the
eliminated-variable stack (the reference's MODEL::extend,
src/gpu/model.cpp:101-162) and verified against the original formula; when
the output is refuted, the input must be refuted as well.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
SAT, UNSAT, UNKNOWN = 10, 20, 0


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libcdcl.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(path)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.cdcl_solve.restype = ctypes.c_int
        L.cdcl_solve.argtypes = [u32, u64, vp, vp, vp, u64, vp, u32, u64, vp]
        L.cdcl_extend.restype = u64
        L.cdcl_extend.argtypes = [vp, u64, vp]
        L.cdcl_verify.restype = u64
        L.cdcl_verify.argtypes = [u64, vp, vp, vp, u64, vp, vp]
        _LIB = L
    return _LIB


def _arr(a, dt):
    return np.ascontiguousarray(np.asarray(a, dtype=dt))


def solve(maxvar, offsets, lits, units=(), assumptions=(), conflict_limit=0):
    """(status, model): status SAT / UNSAT / UNKNOWN; model[v] in {0, 1}
    for v in 0..maxvar (None unless SAT)."""
    off = _arr(offsets, np.uint64)
    li = _arr(lits, np.uint32)
    un = _arr(units, np.uint32)
    asm = _arr(assumptions, np.uint32)
    model = np.zeros(maxvar + 1, np.uint8)
    r = lib().cdcl_solve(maxvar, off.shape[0] - 1, off.ctypes.data, li.ctypes.data, un.ctypes.data, un.shape[0],
                         asm.ctypes.data, asm.shape[0], conflict_limit, model.ctypes.data)
    if r < 0:
        raise MemoryError("cdcl: out of memory")
    return r, (model if r == SAT else None)


def extend(values, stack):
    """MODEL::extend over the stack; values (uint8, 0..maxvar) updated in
    place.  Returns the number of witnesses set."""
    st = _arr(stack, np.uint32)
    n = lib().cdcl_extend(st.ctypes.data, st.shape[0], values.ctypes.data)
    if n == 2**64 - 1:
        raise ValueError("malformed eliminated-variable stack")
    return n


def verify(offsets, lits, values, units=()):
    """(number of falsified clauses and units, index of the first)."""
    off = _arr(offsets, np.uint64)
    li = _arr(lits, np.uint32)
    un = _arr(units, np.uint32)
    first = ctypes.c_uint64(0)
    bad = lib().cdcl_verify(off.shape[0] - 1, off.ctypes.data, li.ctypes.data, un.ctypes.data, un.shape[0],
                            values.ctypes.data, ctypes.byref(first))
    return bad, first.value


def offsets_of(sizes):
    off = np.zeros(len(sizes) + 1, np.uint64)
    np.cumsum(np.asarray(sizes, np.uint64), out=off[1:])
    return off


def original_clauses(sizes, flags, lits):
    """(offsets, literals) of the clauses whose learnt flag (bit 0) is
    clear: the formula a simplified output must stay equisatisfiable with
    (learnt clauses are redundant by contract)."""
    sizes = np.asarray(sizes, np.uint64)
    keep = (np.asarray(flags) & 1) == 0
    mask = np.repeat(keep, sizes.astype(np.int64))
    return offsets_of(sizes[keep]), np.asarray(lits, np.uint32)[mask]


def check_equisat(maxvar, in_offsets, in_lits, in_units, out_status, out_sizes, out_lits, out_units, stack,
                  conflict_limit=2_000_000, input_limit=2_000_000):
    """Checks one simplify call: input (clauses, root units) against the
    output (status 0 UNSAT / 1 SAT / 2 UNSOLVED, clauses, new root units,
    eliminated-variable stack).  Returns a short description of the
    outcome; raises AssertionError on a mismatch and TimeoutError when the
    solver hits its conflict limit."""
    all_units = np.concatenate([_arr(in_units, np.uint32), _arr(out_units, np.uint32)])
    if out_status == 0:
        st, _ = solve(maxvar, in_offsets, in_lits, in_units, conflict_limit=input_limit)
        if st == UNKNOWN:
            raise TimeoutError("input not decided within the conflict limit")
        assert st == UNSAT, "the simplifier refuted a satisfiable input"
        return "unsat (input refuted too)"
    if out_status == 1:
        out_off, out_l = np.zeros(1, np.uint64), np.zeros(0, np.uint32)
    else:
        out_off, out_l = offsets_of(out_sizes), out_lits
    st, model = solve(maxvar, out_off, out_l, all_units, conflict_limit=conflict_limit)
    if st == UNKNOWN:
        raise TimeoutError("output not decided within the conflict limit")
    if st == UNSAT:
        sti, _ = solve(maxvar, in_offsets, in_lits, in_units, conflict_limit=input_limit)
        if sti == UNKNOWN:
            raise TimeoutError("input not decided within the conflict limit")
        assert sti == UNSAT, "the output is unsatisfiable but the input is satisfiable"
        return "unsat (output and input)"
    values = model.copy()
    extend(values, stack)
    bad, first = verify(in_offsets, in_lits, values, in_units)
    assert bad == 0, f"extended model falsifies {bad} input clauses/units (first {first})"
    return "sat (extended model satisfies the input)"
