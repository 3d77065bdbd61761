"""Seeded synthetic CNF workloads (host C++ generators in libpfgen.so).

Formulas are returned as ``Cnf`` objects holding numpy arrays in the
simplifier encoding (literal = 2*v + sign, variables 1..maxvar).  The workload
families are the uniform random k-SAT, the Tseitin-encoded random circuit,
the heavy-tailed k-SAT, batches of the latter two (``batch``), and BMC-style
unrollings (``parafrost_amd.bmc``).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_LIB = None
BASE_SEED = 0x5EED0000


class _PfgCnf(ctypes.Structure):
    _fields_ = [
        ("maxvar", ctypes.c_uint32),
        ("nclauses", ctypes.c_uint64),
        ("nlits", ctypes.c_uint64),
        ("offsets", ctypes.POINTER(ctypes.c_uint64)),
        ("lits", ctypes.POINTER(ctypes.c_uint32)),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpfgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"libpfgen.so not built ({path}); run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.POINTER(_PfgCnf)
        lib.pfg_uniform_ksat.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, P]
        lib.pfg_tseitin.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
        lib.pfg_heavytail.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_uint32, P]
        lib.pfg_prepare.argtypes = [P, P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.pfg_write_dimacs.argtypes = [P, ctypes.c_char_p]
        lib.pfg_free.argtypes = [P]
        _LIB = lib
    return _LIB


@dataclass
class Cnf:
    """A clause list in the simplifier encoding."""

    maxvar: int
    offsets: np.ndarray  # uint64, nclauses + 1
    lits: np.ndarray  # uint32
    flags: np.ndarray | None = None  # uint8 per clause: bit0 learnt, bits 4-5 usage
    lbd: np.ndarray | None = None  # uint32 per clause
    trail: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))  # root units

    @property
    def nclauses(self) -> int:
        return int(self.offsets.shape[0] - 1)

    @property
    def nlits(self) -> int:
        return int(self.lits.shape[0])

    def clause(self, i: int) -> np.ndarray:
        return self.lits[self.offsets[i]:self.offsets[i + 1]]


def _take(c: _PfgCnf) -> Cnf:
    n = int(c.nclauses)
    off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
    lits = np.ctypeslib.as_array(c.lits, shape=(max(int(c.nlits), 1),))[: int(c.nlits)].copy()
    out = Cnf(int(c.maxvar), off, lits)
    _lib().pfg_free(ctypes.byref(c))
    return out


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"{what} failed with code {rc}")


def uniform_ksat(nvars: int, nclauses: int, k: int = 3, seed: int = BASE_SEED + 1) -> Cnf:
    c = _PfgCnf()
    _check(_lib().pfg_uniform_ksat(seed, nvars, nclauses, k, ctypes.byref(c)), "uniform_ksat")
    return _take(c)


def tseitin(ngates: int, ninputs: int = 5000, noutputs: int = 1000, seed: int = BASE_SEED + 2,
            pand: float = 0.6, pxor: float = 0.3, sat_outputs: bool = False) -> Cnf:
    """Tseitin-encoded random circuit; sat_outputs asserts each output with its
    value under a random simulation (satisfiable by construction) instead of
    a random sign."""
    c = _PfgCnf()
    _check(_lib().pfg_tseitin(seed, ninputs, ngates, noutputs, pand, pxor, int(sat_outputs), ctypes.byref(c)),
           "tseitin")
    return _take(c)


def heavytail(nvars: int, nclauses: int, alpha: float = 2.0 / 3.0, maxlen: int = 30,
              seed: int = BASE_SEED + 3) -> Cnf:
    c = _PfgCnf()
    _check(_lib().pfg_heavytail(seed, nvars, nclauses, alpha, maxlen, ctypes.byref(c)), "heavytail")
    return _take(c)


def prepare(cnf: Cnf) -> tuple[Cnf, bool]:
    """Root-level unit propagation and cleanup, as a host solver does before
    calling the simplifier.  Returns (prepared formula, refuted)."""
    src = _PfgCnf(cnf.maxvar, cnf.nclauses, cnf.nlits,
                  cnf.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                  cnf.lits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    dst = _PfgCnf()
    trail = np.zeros(cnf.maxvar + 1, np.uint32)
    ntrail = ctypes.c_uint32(0)
    rc = _lib().pfg_prepare(ctypes.byref(src), ctypes.byref(dst), None,
                            trail.ctypes.data, ctypes.addressof(ntrail))
    out = _take(dst)
    out.trail = trail[: ntrail.value].copy()
    return out, rc == 1


def write_dimacs(cnf: Cnf, path: str) -> None:
    src = _PfgCnf(cnf.maxvar, cnf.nclauses, cnf.nlits,
                  cnf.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                  cnf.lits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    _check(_lib().pfg_write_dimacs(ctypes.byref(src), path.encode()), "write_dimacs")


def from_clauses(maxvar: int, clauses) -> Cnf:
    """Build a Cnf from a list of DIMACS-style signed integer clauses."""
    off = [0]
    lits = []
    for c in clauses:
        for x in c:
            lits.append(2 * abs(x) + (1 if x < 0 else 0))
        off.append(len(lits))
    return Cnf(maxvar, np.array(off, np.uint64), np.array(lits, np.uint32))


def equivalences(nvars: int, nclauses: int, nclasses: int, size: int, k: int = 3, seed: int = BASE_SEED + 5,
                 chords: float = 0.5, units: int = 0, conflict: bool = False, anti: int = 0) -> Cnf:
    """Random k-SAT with planted equivalent literals: nclasses classes of
    'size' literals (distinct variables, random signs), each closed into a
    cycle of implications by binary clauses (-a | b), plus chords; units
    unit clauses of random literals; conflict puts a literal and its
    negation into one class; anti clauses (a | b) or (-a | -b) over two
    variables of one class, each in its own class (they become units or
    tautologies once the class is substituted).  Clause order is shuffled.
    Vectorised: millions of clauses take seconds."""
    rng = np.random.default_rng(seed)
    if nclasses * size > nvars or size < 2:
        raise ValueError("classes need more variables than there are")
    perm = rng.permutation(np.arange(1, nvars + 1, dtype=np.int64))
    members = perm[: nclasses * size].reshape(nclasses, size)
    lits = 2 * members + rng.integers(0, 2, (nclasses, size))
    order = np.argsort(rng.random((nclasses, size)), axis=1)
    cyc = np.take_along_axis(lits, order, axis=1)
    parts = [np.stack([cyc ^ 1, np.roll(cyc, -1, axis=1)], axis=2).reshape(-1, 2)]  # (-a | next a)
    nch = int(chords * size)
    if nch:
        ci = rng.integers(0, size, (nclasses, nch))
        cj = (ci + rng.integers(1, size, (nclasses, nch))) % size
        parts.append(np.stack([np.take_along_axis(lits, ci, axis=1) ^ 1, np.take_along_axis(lits, cj, axis=1)],
                              axis=2).reshape(-1, 2))
    if conflict:
        parts.append(np.array([[lits[0, 0] ^ 1, lits[0, 1] ^ 1], [lits[0, 1], lits[0, 0]]], np.int64))
    na = min(anti, nclasses)
    if na:
        cl = rng.choice(nclasses, na, replace=False)
        ai = rng.integers(0, size, na)
        aj = (ai + rng.integers(1, size, na)) % size
        neg = rng.integers(0, 2, na)
        parts.append(np.stack([2 * members[cl, ai] ^ neg, 2 * members[cl, aj] ^ neg], axis=1))
    bins = np.concatenate(parts)
    # k distinct variables per clause (rows with a repeat are drawn again)
    vs = rng.integers(1, nvars + 1, (nclauses, k))
    while True:
        srt = np.sort(vs, axis=1)
        bad = np.nonzero((srt[:, 1:] == srt[:, :-1]).any(axis=1))[0] if k > 1 else np.zeros(0, np.int64)
        if not bad.shape[0]:
            break
        vs[bad] = rng.integers(1, nvars + 1, (bad.shape[0], k))
    kcl = 2 * vs + rng.integers(0, 2, (nclauses, k))
    ucl = (2 * rng.integers(1, nvars + 1, units) + rng.integers(0, 2, units)).reshape(-1, 1)
    flat = np.concatenate([bins.ravel(), kcl.ravel(), ucl.ravel()]).astype(np.uint32)
    sizes = np.concatenate([np.full(bins.shape[0], 2), np.full(nclauses, k), np.ones(units, np.int64)])
    starts = np.zeros(sizes.shape[0], np.int64)
    np.cumsum(sizes[:-1], out=starts[1:])
    p = rng.permutation(sizes.shape[0])
    ns = sizes[p]
    off = np.zeros(ns.shape[0] + 1, np.uint64)
    np.cumsum(ns, out=off[1:])
    idx = np.repeat(starts[p], ns) + (np.arange(int(off[-1])) - np.repeat(off[:-1].astype(np.int64), ns))
    return Cnf(nvars, off, flat[idx])


def batch(n: int = 512, seed: int = BASE_SEED + 4, cmin: float = 1e3, cmax: float = 5e5) -> list[Cnf]:
    """The batched workload: n formulas with clause counts log-uniform in
    [cmin, cmax]; each is, with probability 1/2, a Tseitin circuit scaled to
    that many clauses (60/30/10% AND/XOR/ITE, satisfiable outputs,
    root-propagated) or a heavy-tailed k-SAT formula with C/4 variables.
    Formula i uses seed ^ i."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(cmin), np.log(cmax), n))).astype(np.int64)
    kinds = rng.integers(0, 2, n)
    out = []
    for i in range(n):
        c = int(sizes[i])
        s = (seed ^ i) & 0xFFFFFFFFFFFFFFFF
        if kinds[i] == 0:
            g = max(100, int(c / 3.4))
            raw = tseitin(g, max(10, g // 400), max(1, g // 2000), s, sat_outputs=True)
            cnf, unsat = prepare(raw)
            if unsat:
                raise RuntimeError(f"batch formula {i} refuted at the root")
            out.append(cnf)
        else:
            out.append(heavytail(max(100, c // 4), c, seed=s))
    return out


# workload definitions -------------------------------------------------------
def workload(name: str, scale: float = 1.0, seed: int | None = None) -> Cnf:
    """The named workloads at full size (scale=1) or scaled down.

    uniform3  : uniform random 3-SAT, 20,000 vars, 85,200 clauses
    tseitin   : 2M-gate random circuit (60% AND, 30% XOR, 10% ITE), 5,000
                inputs, 1,000 outputs asserted with their values under a random
                simulation (satisfiable), root-propagated
    heavytail : 10M vars, 40M clauses, length 2+Geom(1/2) capped at 30,
                variable weights rank^(-2/3)
    """
    if name == "uniform3":
        n = max(10, int(20000 * scale))
        return uniform_ksat(n, int(round(n * 4.26)), 3, seed if seed is not None else BASE_SEED + 1)
    if name == "tseitin":
        g = max(100, int(2_000_000 * scale))
        ni = max(10, int(5000 * min(1.0, scale * 10)))
        no = max(1, min(g, int(1000 * scale)))
        raw = tseitin(g, ni, no, seed if seed is not None else BASE_SEED + 2, sat_outputs=True)
        cnf, unsat = prepare(raw)
        if unsat:
            raise RuntimeError("tseitin workload refuted at the root")
        return cnf
    if name == "heavytail":
        v = max(100, int(10_000_000 * scale))
        return heavytail(v, 4 * v, seed=seed if seed is not None else BASE_SEED + 3)
    raise KeyError(name)
