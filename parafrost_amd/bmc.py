"""BMC-style incremental workload for the serving mode.

A seeded random sequential circuit (default 64 latches, 2,000 gates: 2/3
2-input AND, 1/3 2-input XOR, 32 primary inputs, 8 bad-state outputs) is
Tseitin-encoded frame by frame.  Frame 0 fixes every latch to 0 (unit
clauses); the latches of frame t+1 are the next-state literals of frame t,
so consecutive frames share no fresh variables beyond those literals.  A
frame adds 32 + 2,000 variables and about 6.7k clauses.

A session is one circuit unrolled in requests: each request appends the
clauses of the next frames and names the variables that must stay
(assumed): the latch literals of the newest frame (the next frames connect
there) and 1-32 bad-state outputs of recent frames (the request's
assumption literals).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Request:
    maxvar: int
    offsets: np.ndarray  # uint64, appended clauses
    lits: np.ndarray     # uint32
    assumed: np.ndarray  # uint32 variables kept (frontier latches + assumptions)
    assumptions: np.ndarray  # uint32 literals (bad-state outputs)
    frames: int          # frames in the formula after this request


class Circuit:
    """Random transition circuit over nodes: inputs, current latches, gates."""

    def __init__(self, seed: int, latches: int = 64, gates: int = 2000, inputs: int = 32, bad: int = 8):
        rng = np.random.default_rng(seed)
        self.ni, self.nl, self.ng = inputs, latches, gates
        base = inputs + latches
        self.is_xor = rng.random(gates) < 1.0 / 3.0
        self.fan = np.zeros((gates, 2), np.int64)
        self.fsign = rng.integers(0, 2, (gates, 2))
        for g in range(gates):
            n = base + g  # nodes before this gate
            a, b = rng.choice(n, 2, replace=False)
            self.fan[g] = (a, b)
        # next state of each latch: a gate output with a sign
        self.next = base + rng.choice(gates, latches, replace=gates < latches)
        self.next_sign = rng.integers(0, 2, latches)
        self.bad = base + rng.choice(gates, bad, replace=False)
        self.bad_sign = rng.integers(0, 2, bad)


class Unrolling:
    """Incremental unrolling of a Circuit (variables numbered from 1)."""

    def __init__(self, circuit: Circuit):
        self.c = circuit
        self.maxvar = 0
        self.frames = 0
        self.latch_lits = None  # literals (2v+s) of the current frame's latches
        self.bad_lits: list[np.ndarray] = []

    def _new(self, n):
        v = np.arange(self.maxvar + 1, self.maxvar + 1 + n, dtype=np.int64)
        self.maxvar += n
        return v

    def add_frames(self, k: int):
        """Clauses of the next k frames: (offsets, lits)."""
        c = self.c
        cls: list[np.ndarray] = []
        sizes: list[np.ndarray] = []
        if self.latch_lits is None:
            lv = self._new(c.nl)
            self.latch_lits = 2 * lv
            # initial state: every latch is 0 (unit clauses -latch)
            cls.append((2 * lv + 1).astype(np.uint32))
            sizes.append(np.ones(c.nl, np.int64))
        for _ in range(k):
            iv = self._new(c.ni)
            gv = self._new(c.ng)
            node = np.concatenate([2 * iv, self.latch_lits, 2 * gv])  # node -> literal
            a = node[c.fan[:, 0]] ^ c.fsign[:, 0]
            b = node[c.fan[:, 1]] ^ c.fsign[:, 1]
            o = 2 * gv
            no = o ^ 1
            andm = ~c.is_xor
            # AND o = a & b: (-o a) (-o b) (o -a -b)
            ao, aa, ab, ano = o[andm], a[andm], b[andm], no[andm]
            cls.append(np.stack([ano, aa], 1).ravel())
            sizes.append(np.full(ao.shape[0], 2))
            cls.append(np.stack([ano, ab], 1).ravel())
            sizes.append(np.full(ao.shape[0], 2))
            cls.append(np.stack([ao, aa ^ 1, ab ^ 1], 1).ravel())
            sizes.append(np.full(ao.shape[0], 3))
            # XOR o = a ^ b: (-o a b) (-o -a -b) (o -a b) (o a -b)
            xm = c.is_xor
            xo, xa, xb, xno = o[xm], a[xm], b[xm], no[xm]
            for t in (np.stack([xno, xa, xb], 1), np.stack([xno, xa ^ 1, xb ^ 1], 1),
                      np.stack([xo, xa ^ 1, xb], 1), np.stack([xo, xa, xb ^ 1], 1)):
                cls.append(t.ravel())
                sizes.append(np.full(xo.shape[0], 3))
            self.latch_lits = node[c.next] ^ c.next_sign
            self.bad_lits.append(node[c.bad] ^ c.bad_sign)
            self.frames += 1
        lits = np.concatenate(cls).astype(np.uint32)
        sz = np.concatenate(sizes)
        off = np.zeros(sz.shape[0] + 1, np.uint64)
        np.cumsum(sz, out=off[1:])
        return off, lits

    def request(self, k: int, rng: np.random.Generator) -> Request:
        off, lits = self.add_frames(k)
        recent = np.concatenate(self.bad_lits[-4:])
        m = int(rng.integers(1, min(32, recent.shape[0]) + 1))
        assumptions = rng.choice(recent, m, replace=False).astype(np.uint32)
        assumed = np.unique(np.concatenate([self.latch_lits >> 1, assumptions >> 1])).astype(np.uint32)
        return Request(self.maxvar, off, lits, assumed, assumptions, self.frames)


def session_requests(seed: int, depth: int, step: int = 10, **circuit_kw) -> list[Request]:
    """All requests of one session: the circuit unrolled to 'depth' frames,
    'step' frames per request."""
    rng = np.random.default_rng(seed ^ 0xB3C)
    u = Unrolling(Circuit(seed, **circuit_kw))
    out = []
    while u.frames < depth:
        out.append(u.request(min(step, depth - u.frames), rng))
    return out


def workload(nsessions: int = 16, seed: int = 0x5EED0005, dmin: int = 10, dmax: int = 200, step: int = 10,
             **circuit_kw) -> list[list[Request]]:
    """Sessions with depths uniform in [dmin, dmax]."""
    rng = np.random.default_rng(seed)
    depths = rng.integers(dmin, dmax + 1, nsessions)
    return [session_requests(seed + 17 * i, int(depths[i]), step, **circuit_kw) for i in range(nsessions)]


def poisson_schedule(nreq: int, rate: float, seed: int) -> np.ndarray:
    """Release times (ms) of nreq Poisson arrivals at 'rate' per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1000.0 / rate, nreq))
