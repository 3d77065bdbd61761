"""Benchmark of the MI355X simplifier: one full inprocessing round.

A step is one simplify call with phases=1 and BCE on, i.e. one round of
SUB + BVE (with gate detection) + BCE followed by the stopping round's ERE
(src/cpu/simplify.cpp:184-197), on a formula already resident in HBM: the
device copy of the input is restored (pfs_rewind, device-to-device) and
simplified inside the timed region.  The default workload is the
Tseitin-encoded random circuit with 2M gates (60% AND, 30% XOR, 10% ITE,
5,000 inputs, 1,000 unit outputs, root-propagated: about 6.8M clauses).

Three more legs follow (skip with --no-heavytail / --no-batched / --no-serving):
* heavytail: the 10M-variable heavy-tailed formula (40M clauses), time per
  simplify call;
* batched: 512 formulas (log-uniform 1k-500k clauses, Tseitin or
  heavy-tailed) resident in HBM, simplified as a loop of single-formula
  calls and by the native batch executor; formulas/s of both, per-formula
  p50/p99;
* serving: incremental BMC sessions under Poisson arrivals at 10-500
  requests/s through the native request server; p50/p99 latency and
  throughput per rate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tseitin|uniform3|heavytail]
Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simplify-round time & clauses/s @1 GPU; batched formulas/s + p50/p99 latency"
HBM_PEAK_GBS = 8000.0
OPTIONS = {"phases": 1, "bce": 1}
REF_OPTIONS = ["--phases=1", "-bce"]


def otp_counts(V, C):
    """Range-by-tile counts of the occurrence-table build (otp_plan in
    pfs_kernels.hip): ranges of 4096 literals (8192 past 2048 such ranges),
    tiles of the smallest power of two >= max(4096, 32 x ranges) clauses."""
    n2 = 2 * V + 2
    bits = 12 if (n2 >> 12) < 2048 else 13
    nr = (n2 + (1 << bits) - 1) >> bits
    ct = 4096
    while ct < 32 * nr:
        ct <<= 1
    return nr * ((C + ct - 1) // ct)


def kernel_bytes(name, rounds, cnf):
    """Algorithmic bytes per launch of a kernel, from the work measures of
    the profiled step.  A clause costs a 12-byte header (offset, size,
    status/signature) plus 4 bytes per literal; a list entry costs 4 bytes.

    clause-parallel kernels, per round (C, L, V at round start):
      hist_k     D + 8V            (read every clause, write the 2V histogram)
      scatter_k  D + 4L + 8V       (read every clause and the list offsets, write one ref per literal)
      count_k    D                 (read every clause)
    occurrence table by literal ranges (N = range-by-tile counts, as the
    library plans them: otp_plan):
      otp_count_k         D + 4N   (read every clause, write the counts)
      otp_part_k<true>    D + 4N + 8L   (read every clause and the scanned counts, write the pairs)
      otp_part_k<false>   D + 4N + 4L   (the same with literals only)
      otp_range_k<true>   12L + 24V     (read the pairs, write one ref per literal, offsets, sizes, counts)
      otp_range_k<false>  4L + 8V       (read the literals, write the counts)
    per-elected-variable kernels (O_E list entries, D_E clause bytes of the
    elected variables' lists; elected variables share no clause):
      sub_k, ve_decide_k, bce_k   4 O_E + D_E
      ve_apply_k                  4 O_E + D_E + 4 E (decisions)
    """
    V = cnf.maxvar
    per = []
    for r in rounds:
        c, l = r["clauses_before"], r["literals_before"]
        d = 12 * c + 4 * l
        oe, de = r["elected_occurs"], r["elected_bytes"]
        n = otp_counts(V, c)
        if name == "otp_count_k":
            per.append(d + 4 * n)
        elif name == "otp_part_k<true>":
            per.append(d + 4 * n + 8 * l)
        elif name == "otp_part_k<false>":
            per.append(d + 4 * n + 4 * l)
        elif name == "otp_range_k<true>":
            per.append(12 * l + 24 * V)
        elif name == "otp_range_k<false>":
            per.append(4 * l + 8 * V)
        elif name == "hist_k":
            per.append(d + 8 * V)
        elif name == "scatter_k":
            per.append(d + 4 * l + 8 * V)
        elif name == "count_k":
            per.append(d)
        elif name in ("sub_k", "ve_decide_k", "bce_k"):
            if oe and r["clauses_after"]:  # rounds that ran SUB/VE/BCE (not the stopping round)
                per.append(4 * oe + de)
        elif name == "ve_apply_k":
            if oe and r["clauses_after"]:
                per.append(4 * oe + de + 4 * r["elected"])
        else:
            return None
    if not per or not all(per):
        return None
    return sum(per) / len(per)


def counter_traffic(fetch_csv, write_csv, kernel, copy_bytes=256 << 20):
    """HBM bytes per launch of 'kernel' from two rocprofv3 --pmc runs of this
    script with --calibrate-copy (FETCH_SIZE, then WRITE_SIZE; KiB per
    dispatch).  Each counter is scaled by what it reads for the 256 MiB
    device copy of --calibrate-copy (the three largest copy dispatches).
    Returns (bytes per launch, {"fetch_scale": f, "write_scale": w})."""
    import csv
    import statistics

    def per_kernel(path, counter):
        vals, copies = [], []
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"]
                v = float(r["Counter_Value"]) * 1024.0
                if "copyBuffer" in name or "copy" in name.lower() and "pfs::" not in name:
                    copies.append(v)
                if name.startswith((f"pfs::{kernel}(", f"void pfs::{kernel}(")):  # templates print "void"
                    vals.append(v)
        copies.sort()
        ref = statistics.median(copies[-3:]) if copies else 0.0
        return (sum(vals) / len(vals) if vals else None), (copy_bytes / ref if ref else None)

    fv, fs = per_kernel(fetch_csv, "FETCH_SIZE")
    wv, ws = per_kernel(write_csv, "WRITE_SIZE")
    if fv is None or wv is None or not fs or not ws:
        return None, None
    return fv * fs + wv * ws, {"fetch_scale": round(fs, 4), "write_scale": round(ws, 4)}


def calibrate_copy(nbytes=256 << 20, reps=3):
    """Plain device-to-device copies of a known size, so that hardware
    counter runs over this script can scale FETCH_SIZE/WRITE_SIZE."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").fill_(1)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"bench: calibration copy {nbytes} bytes x {reps}", file=sys.stderr, flush=True)


def with_units(cnf):
    """The prepared formula with its root units as unit clauses: the text
    a host solver would read (its parser and root propagation give back the
    prepared formula)."""
    import numpy as np
    from parafrost_amd import gen
    units = np.asarray(cnf.trail, np.uint32)
    if not units.shape[0]:
        return cnf
    lits = np.concatenate([cnf.lits, units])
    off = np.concatenate([cnf.offsets, cnf.offsets[-1] + 1 + np.arange(units.shape[0], dtype=np.uint64)])
    return gen.Cnf(cnf.maxvar, off, lits)


def bench_ingest(cnf, steps=3):
    """DIMACS ingestion (pfs_parse_dimacs) of the workload's text, the text
    the CPU baseline's reference parses.  Wall time includes the host-to-
    device copy of the text; the emit lexer's roofline counts the text read
    once and the literal stream written (N + 4 T bytes, T tokens)."""
    import numpy as np
    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier
    full = with_units(cnf)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.cnf")
        gen.write_dimacs(full, p)
        with open(p, "rb") as f:
            text = f.read()
    s = Simplifier(0)
    try:
        s.parse_dimacs(text)
        t0 = time.perf_counter()
        for _ in range(steps):
            info = s.parse_dimacs(text)
        dt = (time.perf_counter() - t0) / steps
        s.set_option("profilegpu", 1)
        s.parse_dimacs(text)
        kt = s.kernel_times()
        s.set_option("profilegpu", 0)
        r = s.result(2)
        trail = s.trail()
    finally:
        s.close()
    sizes = np.diff(cnf.offsets).astype(np.int64)
    order = np.lexsort((cnf.lits, np.repeat(np.arange(cnf.nclauses), sizes)))
    same = (np.array_equal(r.sizes, sizes) and np.array_equal(r.lits, cnf.lits[order])
            and np.array_equal(np.sort(trail), np.sort(np.asarray(cnf.trail, np.uint32))))
    T = full.nlits + full.nclauses
    emit = kt.get("lex_k<LEX_EMIT>")
    out = {"text_bytes": len(text), "ms": round(dt * 1e3, 3), "text_gbs_pcie_inclusive": round(len(text) / dt / 1e9, 2),
           "clauses_per_s": round(cnf.nclauses / dt, 1), "device_kernel_ms": round(sum(v[0] for v in kt.values()), 3),
           "kernels_ms": {k: round(v[0], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:8]},
           "identical_to_prepared": bool(same), "read_clauses": info["read_clauses"]}
    if emit and emit[1]:
        us = emit[0] / emit[1] * 1e3
        ach = (len(text) + 4 * T) / (us * 1e-6) / 1e9
        out["lex_emit"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(us, 2),
                           "algorithmic_bytes_per_launch": len(text) + 4 * T}
    print("bench: ingest " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def bench_els(steps=3, rounds=2):
    """Equivalent-literal substitution (pfs_decompose) on a planted-
    equivalence formula: 2,000,000 variables, 6,000,000 ternary clauses,
    200,000 classes of 5 literals closed by binary clauses (+ chords).
    Inputs resident in HBM (pfs_rewind between calls).  The reference's own
    decompose rounds on the same text (oracle/_ref/pfref, PFREF_ELS_OUT, one
    core) are timed beside it."""
    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier
    cnf = gen.equivalences(2_000_000, 6_000_000, 200_000, 5, seed=0x5EED0007, chords=0.4)
    s = Simplifier(0, keep_input=True, decomposemin=rounds)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits)
        s.decompose()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.rewind()
            info = s.decompose()
        dt = (time.perf_counter() - t0) / steps
        s.set_option("profilegpu", 1)
        s.rewind()
        s.decompose()
        kt = s.kernel_times()
        s.set_option("profilegpu", 0)
    finally:
        s.close()
    out = {"variables": cnf.maxvar, "clauses": cnf.nclauses, "rounds": info["rounds"],
           "substituted": info["substituted"], "ms": round(dt * 1e3, 3), "clauses_per_s": round(cnf.nclauses / dt, 1),
           "device_kernel_ms": round(sum(v[0] for v in kt.values()), 3),
           "kernels_ms": {k: round(v[0], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:8]}}
    ref = os.path.join(ROOT, "oracle", "_ref", "pfref")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "e.cnf")
            gen.write_dimacs(cnf, p)
            env = dict(os.environ, PFREF_ELS_OUT=os.path.join(d, "e.els"), PFREF_ELS_ROUNDS=str(rounds))
            r = subprocess.run(["taskset", "-c", "0", ref, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd")],
                               capture_output=True, text=True, timeout=600, env=env)
            line = [x for x in r.stdout.splitlines() if x.startswith("{")]
            if line:
                j = json.loads(line[-1])
                out["cpu_reference_s"] = j.get("els_s")
                out["cpu_reference_rounds"] = j.get("els_rounds")
    print("bench: els " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def cpu_baseline(cnf, name, budget_clauses=7_500_000):
    """The reference CPU simplifier (compiled from its sources into
    oracle/_ref/pfref) on the same workload, or a bounded sample of it, on one
    host core; falls back to the oracle restatement when the reference
    binary is absent."""
    from parafrost_amd import gen
    sample = cnf
    note = "full workload"
    if cnf.nclauses > budget_clauses:
        scale = budget_clauses / cnf.nclauses
        sample = gen.workload(name, scale=scale)
        note = f"{name} workload scaled by {scale:.3f} ({sample.nclauses} clauses)"
    ref = os.path.join(ROOT, "oracle", "_ref", "pfref")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "w.cnf")
            # the reference runs its own parser and root propagation on the
            # raw formula; the prepared formula has the same clauses and units
            gen.write_dimacs(with_units(sample), p)
            r = subprocess.run(["taskset", "-c", "0", ref, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd")]
                               + REF_OPTIONS, capture_output=True, text=True, timeout=600)
            line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
            t = json.loads(line)["total_s"]
            tp = json.loads(line).get("parse_s")
        return {"value": sample.nclauses / t, "unit": "clauses/s", "cores": 1, "kind": "reference",
                "sample": note + f"; reference CPU simplifier (ParaFROST src/cpu) round time {t:.3f} s, taskset 1 core",
                "seconds": t, "parse_seconds": tp, "parse_sample": note}
    from oracle import oracle
    r = oracle.simplify(sample.maxvar, sample.offsets, sample.lits, nassigned=int(sample.trail.shape[0]), det=0,
                        phases=1, bce_en=1)
    return {"value": sample.nclauses / r.seconds, "unit": "clauses/s", "cores": 1, "kind": "port",
            "sample": note + f"; oracle restatement round time {r.seconds:.3f} s", "seconds": r.seconds}


def bench_single(name, steps=3, warmup=1):
    """Another single-formula workload (default: the 10M-variable heavy-tailed
    formula, 40M clauses): ms per simplify call with inputs resident."""
    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier
    t0 = time.time()
    cnf = gen.workload(name)
    s = Simplifier(0, keep_input=True, **OPTIONS)
    s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
    print(f"bench: {name}: {cnf.maxvar} vars, {cnf.nclauses} clauses ({time.time() - t0:.1f}s to generate "
          "and upload)", file=sys.stderr, flush=True)
    status = None
    for _ in range(warmup):
        s.rewind()
        status = s.simplify()
    t = time.perf_counter()
    for _ in range(steps):
        s.rewind()
        status = s.simplify()
    dt = (time.perf_counter() - t) / steps
    r = s.result(status).rounds
    s.close()
    out = {"workload": name, "variables": cnf.maxvar, "clauses": cnf.nclauses, "literals": cnf.nlits,
           "ms_per_step": round(dt * 1e3, 3), "clauses_per_s": round(cnf.nclauses / dt, 1),
           "round_result": ["UNSAT", "SAT", "UNSOLVED"][status], "rounds": len(r),
           "sub_passes": [x["sub_passes"] for x in r]}
    print("bench: " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def bench_handoff(s, cnf):
    """The handoff after a simplify call on the same workload: the packed
    download of the result (device packing + PCIe) and the device variable
    map (pfs_map, forced).  The map kernel's roofline counts a read of the
    clause handles and clauses (4C + 4(3C + L)) and a write of the
    literals and signatures (4L + 4C), C and L of the result."""
    import numpy as np
    s.rewind()
    status = s.simplify()
    if status == 0:  # refuted: nothing to hand off (pfs_map refuses a refuted formula)
        return {"skipped": "the formula is refuted"}
    t0 = time.perf_counter()
    res = s.result(status)
    dl_ms = (time.perf_counter() - t0) * 1e3
    C, L = int(res.sizes.shape[0]), int(res.lits.shape[0])
    dl_bytes = 8 * (C + 1) + 4 * L + C + 4 * C + 4 * res.units.shape[0] + 4 * res.resolved.shape[0] + res.vstate.shape[0]
    s.set_option("profilegpu", 1)
    t0 = time.perf_counter()
    info = s.map(force=True)
    map_ms = (time.perf_counter() - t0) * 1e3
    kt = s.kernel_times()
    s.set_option("profilegpu", 0)
    k_ms, k_n = kt.get("map_clauses_k", (0.0, 0))
    alg = 4 * C + 4 * (3 * C + L) + 4 * L + 4 * C
    out = {"download_ms": round(dl_ms, 3), "download_bytes": int(dl_bytes),
           "download_gbs": round(dl_bytes / dl_ms / 1e6, 2) if dl_ms else None,
           "map_ms": round(map_ms, 3), "old_vars": info["old_maxvar"], "new_vars": info["new_maxvar"],
           "clauses": C, "literals": L,
           "map_kernels_ms": {k: round(v[0], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])},
           "map_clauses_k": {"bound": "hbm", "achieved": round(alg / (k_ms / k_n / 1e3) / 1e9, 1) if k_n else None,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(alg / (k_ms / k_n / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if k_n else None,
                             "algorithmic_bytes_per_launch": int(alg),
                             "avg_launch_us": round(k_ms / k_n * 1e3, 2) if k_n else None}}
    # the map's own check: the mapped clauses are the map of the unmapped ones
    m = s.download_map(info["old_maxvar"])
    after = s.result(status)
    expect = (2 * m[res.lits >> 1] + (res.lits & 1)).astype(np.uint32)
    out["mapped_identical"] = bool(np.array_equal(after.lits, expect) and np.array_equal(after.sizes, res.sizes))
    print("bench: handoff " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def bench_batched(n=512, workers=16, check=16):
    """Batched workload: n formulas (clause counts log-uniform in 1k-500k,
    Tseitin or heavy-tailed) resident in HBM, simplified once as a loop of
    plain single-formula calls and once by pfs_batch_simplify with 'workers'
    native threads (one stream per formula).  formulas/s of both, per-formula
    latency percentiles of the batched run, and a bit-exact comparison of
    the first 'check' outputs."""
    import numpy as np
    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier, batch_simplify
    t0 = time.time()
    cnfs = gen.batch(n)
    ncl = sum(c.nclauses for c in cnfs)
    ctx = []
    for c in cnfs:
        s = Simplifier(0, keep_input=True, **OPTIONS)
        s.upload(c.maxvar, c.offsets, c.lits, trail=c.trail)
        ctx.append(s)
    print(f"bench: batch of {n} formulas, {ncl} clauses ({time.time() - t0:.1f}s to generate and upload)",
          file=sys.stderr, flush=True)
    # warm-up (kernel loading, buffer growth), then the timed runs
    batch_simplify(ctx[:workers], workers)
    for s in ctx[:workers]:
        s.rewind()
    t = time.perf_counter()
    loop_res = [s.simplify() for s in ctx]
    loop_s = time.perf_counter() - t
    ref = [ctx[i].result(loop_res[i]) for i in range(min(check, n))]
    for s in ctx:
        s.rewind()
    t = time.perf_counter()
    res, ms = batch_simplify(ctx, workers)
    batch_s = time.perf_counter() - t
    same = 0
    for i in range(min(check, n)):
        b = ctx[i].result(res[i])
        same += int(b.status == ref[i].status and np.array_equal(b.lits, ref[i].lits) and
                    np.array_equal(b.sizes, ref[i].sizes) and np.array_equal(b.resolved, ref[i].resolved) and
                    np.array_equal(b.flags, ref[i].flags) and np.array_equal(b.vstate, ref[i].vstate))
    for s in ctx:
        s.close()
    out = {"formulas": n, "clauses": ncl, "workers": workers,
           "formulas_per_s": round(n / batch_s, 1), "loop_formulas_per_s": round(n / loop_s, 1),
           "speedup_vs_loop": round(loop_s / batch_s, 2), "clauses_per_s": round(ncl / batch_s, 1),
           "p50_ms": round(float(np.percentile(ms, 50)), 3), "p99_ms": round(float(np.percentile(ms, 99)), 3),
           "batch_s": round(batch_s, 3), "loop_s": round(loop_s, 3), "identical_checked": f"{same}/{min(check, n)}"}
    print("bench: batched " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def bench_serving(rates=(10, 20, 50, 100, 200, 500), nsessions=16, nreq=64, workers=8):
    """Serving workload: BMC sessions (64 latches, 2,000 gates, depth
    uniform in 10-200 frames, 10 frames per request, 1-32 assumptions) whose
    formulas stay resident; requests arrive as a Poisson stream at each rate
    and are served by the request server (continuous batching over
    'workers' native threads).  Latency = completion - arrival."""
    import numpy as np
    from parafrost_amd import bmc
    from parafrost_amd.simplifier import Server, Simplifier
    t0 = time.time()
    sessions = bmc.workload(nsessions)
    print(f"bench: serving {nsessions} sessions, {sum(len(x) for x in sessions)} requests available "
          f"({time.time() - t0:.1f}s to generate)", file=sys.stderr, flush=True)
    sweep = []
    for rate in rates:
        rng = np.random.default_rng(rate)
        ctx = []
        for _ in sessions:
            s = Simplifier(0, **OPTIONS)
            s.upload(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
            ctx.append(s)
        pending = [0] * nsessions
        order = []
        while len(order) < nreq:
            live = [i for i in range(nsessions) if pending[i] < len(sessions[i])]
            if not live:
                break
            si = int(rng.choice(live))
            order.append((si, pending[si]))
            pending[si] += 1
        times = bmc.poisson_schedule(len(order), rate, rate + 1)
        srv = Server(workers)
        tickets = []
        for (si, q), rel in zip(order, times):
            rq = sessions[si][q]
            tickets.append(srv.submit(ctx[si], float(rel), rq.maxvar, rq.offsets, rq.lits, rq.assumed))
        srv.start()
        lat, ends = [], []
        for tk in tickets:
            st, res, rel, start, end = srv.wait(tk)
            if st != 0:
                raise RuntimeError(f"serving request failed with status {st}")
            lat.append(end - rel)
            ends.append(end)
        srv.close()
        for s in ctx:
            s.close()
        span = max(ends) - float(times[0])
        row = {"rate_rps": rate, "requests": len(order), "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "throughput_rps": round(len(order) / (span / 1e3), 1)}
        sweep.append(row)
        print("bench: serving " + json.dumps(row), file=sys.stderr, flush=True)
    return {"sessions": nsessions, "workers": workers, "frames_per_request": 10, "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tseitin", choices=["tseitin", "uniform3", "heavytail"])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="extra simplifier option (launch shaping such as workorder=0)")
    ap.add_argument("--traffic-csv", default=None, metavar="FETCH.csv,WRITE.csv",
                    help="rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE counter files of this workload "
                         "(runs with --calibrate-copy) for roofline.traffic")
    ap.add_argument("--no-heavytail", action="store_true", help="skip the heavy-tailed single-formula leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-ingest", action="store_true", help="skip the DIMACS ingestion leg")
    ap.add_argument("--no-els", action="store_true", help="skip the equivalent-literal substitution leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run 256 MiB device copies first (reference kernel for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(formulas shard by independent instances or elected variables; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier

    t0 = time.time()
    cnf = gen.workload(args.workload, scale=args.scale)
    gen_s = time.time() - t0
    print(f"bench: workload {args.workload} scale {args.scale}: {cnf.maxvar} vars, {cnf.nclauses} clauses, "
          f"{cnf.nlits} literals ({gen_s:.1f}s to generate)", file=sys.stderr, flush=True)

    if args.calibrate_copy:
        calibrate_copy()
    opts = dict(OPTIONS)
    for kv in args.set:
        k, v = kv.split("=", 1)
        opts[k] = int(v)
    s = Simplifier(0, keep_input=True, **opts)
    s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    status = None
    for _ in range(args.warmup):
        s.rewind()
        status = s.simplify()
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.rewind()
        status = s.simplify()
    sync()
    dt = (time.perf_counter() - t0) / args.steps
    res = s.result(status)
    print(f"bench: {args.steps} steps, {dt * 1e3:.2f} ms/step, result {res.status}, "
          f"{res.sizes.shape[0]} clauses out, rounds {res.rounds}", file=sys.stderr, flush=True)

    # one profiled step: per-kernel device time from HIP events on the
    # context's stream
    s.set_option("profilegpu", 1)
    s.rewind()
    s.simplify()
    kt = s.kernel_times()
    stages = s.stage_times()
    s.set_option("profilegpu", 0)
    rounds = s.result(status).rounds
    dom = max(kt.items(), key=lambda kv: kv[1][0], default=None)
    print("bench: kernels " + json.dumps({k: [round(v[0], 3), v[1]] for k, v in
                                          sorted(kt.items(), key=lambda kv: -kv[1][0])[:12]}),
          file=sys.stderr, flush=True)
    print("bench: stages " + json.dumps({k: [round(v[0], 3), v[1]] for k, v in stages.items()}),
          file=sys.stderr, flush=True)
    hbm = max(((k, v) for k, v in kt.items() if kernel_bytes(k, rounds, cnf)), key=lambda kv: kv[1][0],
              default=None)
    roof_kernel = dom if dom and kernel_bytes(dom[0], rounds, cnf) else hbm
    roofline = None
    if roof_kernel:
        name, (ms, n) = roof_kernel
        avg_s = ms / n / 1e3
        b = kernel_bytes(name, rounds, cnf)
        ach = b / avg_s / 1e9
        traffic, scales = None, None
        if args.traffic_csv:
            fcsv, wcsv = args.traffic_csv.split(",")
            traffic, scales = counter_traffic(fcsv, wcsv, name)
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": int(traffic) if traffic else None, "traffic_scales": scales,
                    "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes_per_launch": int(b),
                    "dominant_kernel": dom[0], "dominant_kernel_ms_per_step": round(dom[1][0], 3),
                    "modelled_kernels": {
                        k: {"frac": round(kernel_bytes(k, rounds, cnf) / (v[0] / v[1] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                            "avg_launch_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1]}
                        for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0]) if kernel_bytes(k, rounds, cnf)}}
    handoff = bench_handoff(s, cnf)
    s.close()
    ingest = None if args.no_ingest else bench_ingest(cnf)
    els_leg = None if args.no_els else bench_els()

    heavy = None if (args.no_heavytail or args.workload == "heavytail") else bench_single("heavytail")
    batched = None if args.no_batched else bench_batched()
    serving = None if args.no_serving else bench_serving()

    base = None
    if not args.no_cpu_baseline:
        try:
            base = cpu_baseline(cnf, args.workload)
        except Exception as e:  # the baseline is reported, never required
            print(f"bench: cpu baseline failed: {e}", file=sys.stderr)
    if ingest is not None and base and base.get("parse_seconds"):
        ingest["cpu_reference_parse_s"] = base["parse_seconds"]
        ingest["cpu_reference_parse_sample"] = base.get("parse_sample")
    value = cnf.nclauses / dt
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "clauses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} (scale {args.scale})", "variables": cnf.maxvar,
                   "clauses": cnf.nclauses, "literals": cnf.nlits, "options": "phases=1 bce=1 (SUB+BVE+BCE, ERE)",
                   "round_result": ["UNSAT", "SAT", "UNSOLVED"][status], "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": base,
        "handoff": handoff,
        "ingest": ingest,
        "els": els_leg,
        "heavytail": heavy,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
