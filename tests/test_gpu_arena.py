"""The device arena on the GPU (pfs_arena_create / pfs_create_in_arena).

* contexts in a private arena give the same results as in the default one;
* BMC sessions whose resident formulas share a tight private arena, served
  in an interleaved order, force compactions (the idle sessions' dynamic
  blocks are moved while another session grows) and still give every
  request the result it gets alone; the same with eight workers in a roomier
  arena, where busy sessions' blocks stay pinned;
* an arena too small for a formula refuses it with PFS_ALLOC_FAIL.
"""
import numpy as np
import pytest

from parafrost_amd import bmc, gen
from tests.common import diff

pytestmark = pytest.mark.gpu
OPTS = {"phases": 2, "bce": 1}


def test_private_arena_matches_default():
    from parafrost_amd.simplifier import Arena, Simplifier
    cnfs = [gen.uniform_ksat(3000, 12780, 3, seed=1), gen.heavytail(20000, 80000, seed=2),
            gen.prepare(gen.tseitin(20000, 400, 40, seed=3))[0]]
    arena = Arena(512 << 20)
    try:
        for cnf in cnfs:
            res = []
            for a in (None, arena):
                s = Simplifier(0, arena=a, phases=1, bce=1)
                try:
                    s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
                    res.append(s.result(s.simplify()))
                finally:
                    s.close()
            d = diff(res[0], res[1], units_as_set=True)
            assert not d, d
        st = arena.stats()
        assert st["capacity"] == 512 << 20 and st["slabs"] == 1 and st["used"] == 0
    finally:
        arena.close()


def _sessions():
    return [bmc.session_requests(70 + i, depth=int(d), step=10, gates=800, latches=32)
            for i, d in enumerate([40, 60, 30, 50, 20, 45])]


def _run(sessions, schedule, workers=1, arena=None):
    """Serves the schedule (session, request) through the request server;
    returns the result codes per schedule entry, the final states and the
    contexts' (used, peak) bytes.  The server is closed (every submitted
    request finished) before any context, also when a request failed."""
    from parafrost_amd.simplifier import Server, Simplifier
    ctx, srv = [], None
    try:
        for _ in sessions:
            s = Simplifier(0, arena=arena, **OPTS)
            s.upload(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
            ctx.append(s)
        srv = Server(workers, 0)
        tickets = [srv.submit(ctx[si], 0.0, sessions[si][q].maxvar, sessions[si][q].offsets,
                              sessions[si][q].lits, sessions[si][q].assumed) for si, q in schedule]
        srv.start()
        results = [srv.wait(t) for t in tickets]
        srv.close()
        srv = None
        bad = [r[0] for r in results if r[0] != 0]
        assert not bad, f"{len(bad)} requests failed (status {bad[0]})"
        codes = [r[1] for r in results]
        finals = [s.result(2) for s in ctx]
        mem = [s.memory() for s in ctx]
        return codes, finals, mem
    finally:
        if srv is not None:
            srv.close()
        for s in ctx:
            s.close()


def _interleaved(sessions, seed):
    rng = np.random.default_rng(seed)
    pending = [0] * len(sessions)
    out = []
    while any(pending[i] < len(sessions[i]) for i in range(len(sessions))):
        live = [i for i in range(len(sessions)) if pending[i] < len(sessions[i])]
        si = int(rng.choice(live))
        out.append((si, pending[si]))
        pending[si] += 1
    return out


def test_tight_arena_compacts_and_mixed_equals_alone():
    """Interleaved sessions in private arenas from barely above the sum of
    the sessions' peaks upwards: every arena that serves all requests gives
    each session its result alone, and the tightest ones do so by compacting
    (idle sessions' blocks moved while another grows)."""
    from parafrost_amd.simplifier import Arena
    sessions = _sessions()
    alone = [(si, q) for si in range(len(sessions)) for q in range(len(sessions[si]))]
    codes_a, final_a, mem = _run(sessions, alone)
    # a context keeps its buffers at their largest: the sessions end holding
    # the sum of their peaks
    resident = sum(p for _, p in mem)
    order_a = {k: i for i, k in enumerate(alone)}
    mixed = _interleaved(sessions, 11)
    compacted = served = 0
    for slack in (0.02, 0.04, 0.06, 0.07, 0.08, 0.09, 0.10, 0.11, 0.12, 0.25):
        cap = int(resident * (1 + slack)) + (1 << 20)
        arena = Arena(cap)
        try:
            try:
                codes_m, final_m, _ = _run(sessions, mixed, workers=1, arena=arena)
            except AssertionError as e:  # a request refused with PFS_ALLOC_FAIL: too tight
                print("arena", cap, "too small:", e, flush=True)
                continue
        finally:
            st = arena.stats()
            print("arena", cap, "stats", st, flush=True)
            arena.close()
        served += 1
        for i, k in enumerate(mixed):
            assert codes_m[i] == codes_a[order_a[k]]
        for a, m in zip(final_a, final_m):
            d = diff(a, m, units_as_set=True)
            assert not d, d
        compacted += st["compactions"] > 0 and st["moved_bytes"] > 0
    assert served >= 2 and compacted >= 1, (served, compacted)


def test_concurrent_sessions_in_shared_arena():
    from parafrost_amd.simplifier import Arena
    sessions = _sessions()
    alone = [(si, q) for si in range(len(sessions)) for q in range(len(sessions[si]))]
    codes_a, final_a, mem = _run(sessions, alone)
    cap = 2 * sum(p for _, p in mem) + (64 << 20)
    arena = Arena(cap)
    try:
        mixed = _interleaved(sessions, 12)
        codes_m, final_m, _ = _run(sessions, mixed, workers=8, arena=arena)
    finally:
        arena.close()
    order_a = {k: i for i, k in enumerate(alone)}
    for i, k in enumerate(mixed):
        assert codes_m[i] == codes_a[order_a[k]]
    for a, m in zip(final_a, final_m):
        d = diff(a, m, units_as_set=True)  # units of one call come in propagation order
        assert not d, d


def test_arena_too_small_refuses():
    from parafrost_amd.simplifier import Arena, PfsError, Simplifier
    cnf = gen.uniform_ksat(20000, 85200, 3, seed=4)
    arena = Arena(1 << 20)
    try:
        s = Simplifier(0, arena=arena)
        try:
            with pytest.raises(PfsError, match="ALLOC_FAIL"):
                s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        finally:
            s.close()
        assert arena.stats()["used"] == 0
    finally:
        arena.close()
