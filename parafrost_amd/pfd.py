"""PFD1 dump format: one simplifier state (input or output) in a file.

Layout (little endian):
  "PFD1" | u32 status (0 UNSAT, 1 SAT, 2 UNSOLVED) | u32 maxvar
  | u64 ncls | u32 sizes[ncls] | u8 flags[ncls] | u32 lbd[ncls]
  | u64 nlits | u32 lits[nlits]
  | u64 nunits | u32 units[nunits]        (input: root trail; output: new units)
  | u64 nres | u32 resolved[nres]         (eliminated-variable stack)
  | u64 nvst | u8 vstate[nvst]            (per variable, index 0 unused)
Clause flags: bit0 learnt, bit1 added (resolvent), bits 4-5 usage.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass

import numpy as np


@dataclass
class State:
    status: int
    maxvar: int
    sizes: np.ndarray
    flags: np.ndarray
    lbd: np.ndarray
    lits: np.ndarray
    units: np.ndarray
    resolved: np.ndarray
    vstate: np.ndarray

    @property
    def offsets(self) -> np.ndarray:
        off = np.zeros(self.sizes.shape[0] + 1, np.uint64)
        np.cumsum(self.sizes, out=off[1:])
        return off

    def clauses(self):
        off = self.offsets
        return [tuple(int(x) for x in self.lits[off[i]:off[i + 1]]) for i in range(self.sizes.shape[0])]


def read(path: str) -> State:
    """Reads a PFD1 file (gzip-compressed when the name ends in .gz)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        b = f.read()
    if b[:4] != b"PFD1":
        raise ValueError(f"{path}: not a PFD1 file")
    pos = 4

    def take(dt, n):
        nonlocal pos
        a = np.frombuffer(b, dtype=dt, count=n, offset=pos).copy()
        pos += a.nbytes
        return a

    status, maxvar = (int(x) for x in take(np.uint32, 2))
    n = int(take(np.uint64, 1)[0])
    sizes, flags, lbd = take(np.uint32, n), take(np.uint8, n), take(np.uint32, n)
    lits = take(np.uint32, int(take(np.uint64, 1)[0]))
    units = take(np.uint32, int(take(np.uint64, 1)[0]))
    resolved = take(np.uint32, int(take(np.uint64, 1)[0]))
    vstate = take(np.uint8, int(take(np.uint64, 1)[0]))
    return State(status, maxvar, sizes, flags, lbd, lits, units, resolved, vstate)


def write(path: str, s: State) -> None:
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(b"PFD1")
        f.write(np.array([s.status, s.maxvar], np.uint32).tobytes())
        f.write(np.array([s.sizes.shape[0]], np.uint64).tobytes())
        f.write(s.sizes.astype(np.uint32).tobytes())
        f.write(s.flags.astype(np.uint8).tobytes())
        f.write(s.lbd.astype(np.uint32).tobytes())
        for arr, dt in ((s.lits, np.uint32), (s.units, np.uint32), (s.resolved, np.uint32), (s.vstate, np.uint8)):
            f.write(np.array([arr.shape[0]], np.uint64).tobytes())
            f.write(arr.astype(dt).tobytes())
