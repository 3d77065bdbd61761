"""Incremental sessions and the request server.

* every step of an incremental BMC session (pfs_append + pfs_assume +
  pfs_simplify on the resident formula) equals the oracle run on the same
  state: the previous result, the appended clauses after the incremental
  clean-up, the pending units;
* the request server gives every session the same final state alone and
  mixed with other sessions under Poisson arrivals, with several workers
  and under a tight memory budget.
"""
import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import bmc
from tests.common import append_clean, diff, oracle_kw, proof_lines_canonical

pytestmark = pytest.mark.gpu

OPTS = {"phases": 2, "bce": 1}


def _flat(clauses):
    off = np.zeros(len(clauses) + 1, np.uint64)
    if clauses:
        np.cumsum([len(c) for c in clauses], out=off[1:])
    lits = np.array([x for c in clauses for x in c], np.uint32)
    return off, lits


@pytest.mark.parametrize("seed", [1, 2])
def test_incremental_steps_match_oracle(seed):
    from parafrost_amd.simplifier import Simplifier
    reqs = bmc.session_requests(seed, depth=25, step=10, gates=600, latches=24, inputs=16)
    s = Simplifier(0, proof=1, **OPTS)
    s.upload(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    prev_clauses, assign, trail, nres, nunits = [], {}, [], 0, 0
    vstate = np.zeros(1, np.uint8)
    try:
        for q, rq in enumerate(reqs):
            print("incremental", seed, q, rq.maxvar, rq.offsets.shape[0] - 1, flush=True)
            kept, new_units, refuted = append_clean(assign, rq.offsets, rq.lits)
            assert not refuted
            s.append(rq.maxvar, rq.offsets, rq.lits)
            s.assume(rq.assumed)
            g = s.result(s.simplify())
            off, lits = _flat(prev_clauses + kept)
            vs = np.zeros(rq.maxvar + 1, np.uint8)
            vs[: vstate.shape[0]] = vstate
            for u in new_units:
                vs[u >> 1] = 2
            asm = np.zeros(rq.maxvar + 1, np.uint8)
            asm[rq.assumed] = 1
            r = oracle.simplify(rq.maxvar, off, lits, vstate=vs, assumed=asm, trail=trail + new_units,
                                npropagated=len(trail), det=1, proof_en=1, **oracle_kw(OPTS))
            assert g.status == r.status
            if r.status == 0:
                break
            res_step = g.resolved[nres:]
            units_step = set(g.units[nunits:].tolist())
            g_step = type(g)(g.status, g.maxvar, g.sizes, g.flags, g.lbd, g.lits, g.units[nunits:], res_step,
                             g.vstate, g.rounds)
            d = diff(g_step, r)  # units as sets; stack: this step's words
            assert d == ["units differ ({} vs {})".format(len(units_step), len(r.units))] or not d, d
            assert units_step == set(new_units) | set(r.units.tolist())
            assert proof_lines_canonical(g.proof) == proof_lines_canonical(r.proof)
            # next state: the GPU's own result
            o = np.concatenate([[0], np.cumsum(g.sizes.astype(np.int64))])
            prev_clauses = [g.lits[o[i]:o[i + 1]].tolist() for i in range(g.sizes.shape[0])]
            for u in g.units[nunits:].tolist():
                assign[u >> 1] = u
            trail = trail + [u for u in g.units[nunits:].tolist()]
            vstate = g.vstate.copy()
            nres, nunits = g.resolved.shape[0], g.units.shape[0]
    finally:
        s.close()


def _final_states(sessions, workers, schedule, budget=0):
    from parafrost_amd.simplifier import Server, Simplifier
    ctx = []
    for _ in sessions:
        s = Simplifier(0, **OPTS)
        s.upload(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
        ctx.append(s)
    srv = Server(workers, budget)
    tickets = []
    for i, (si, q) in enumerate(schedule):
        rq = sessions[si][q]
        tickets.append(srv.submit(ctx[si], schedule_time(schedule, i), rq.maxvar, rq.offsets, rq.lits, rq.assumed))
    srv.start()
    codes = []
    for t in tickets:
        st, res, rel, start, end = srv.wait(t)
        assert st == 0, st
        assert end >= start >= rel - 1e-6
        codes.append(res)
    srv.close()
    out = []
    for s in ctx:
        out.append(s.result(2))
        s.close()
    return codes, out


_TIMES = {}


def schedule_time(schedule, i):
    return _TIMES.get(id(schedule), [0.0] * len(schedule))[i]


def test_server_mixed_equals_alone():
    sessions = [bmc.session_requests(50 + i, depth=int(d), step=10, gates=800, latches=32)
                for i, d in enumerate([20, 35, 50, 25, 40, 30])]
    # alone: one worker, sessions one after the other
    alone_sched = [(si, q) for si in range(len(sessions)) for q in range(len(sessions[si]))]
    codes_a, final_a = _final_states(sessions, 1, alone_sched)
    # mixed: Poisson arrivals over interleaved sessions, 8 workers
    rng = np.random.default_rng(7)
    pending = [0] * len(sessions)
    mixed = []
    while any(pending[i] < len(sessions[i]) for i in range(len(sessions))):
        live = [i for i in range(len(sessions)) if pending[i] < len(sessions[i])]
        si = int(rng.choice(live))
        mixed.append((si, pending[si]))
        pending[si] += 1
    _TIMES[id(mixed)] = bmc.poisson_schedule(len(mixed), 400.0, 9).tolist()
    codes_m, final_m = _final_states(sessions, 8, mixed)
    order_a = {k: i for i, k in enumerate(alone_sched)}
    for i, k in enumerate(mixed):
        assert codes_m[i] == codes_a[order_a[k]]
    for a, m in zip(final_a, final_m):
        d = diff(a, m, units_as_set=True)
        assert not d, d
    # a memory budget that admits one session at a time
    codes_b, final_b = _final_states(sessions, 8, mixed, budget=1)
    for a, b in zip(final_a, final_b):
        assert not diff(a, b, units_as_set=True)
