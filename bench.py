"""Benchmark of the MI355X simplifier: one full inprocessing round.

A step is one simplify call with phases=1 and BCE on, i.e. one round of
SUB + BVE (with gate detection) + BCE followed by the stopping round's ERE
(src/cpu/simplify.cpp:184-197), on a formula already resident in HBM: the
device copy of the input is restored (pfs_rewind, device-to-device) and
simplified inside the timed region.  The default workload (the headline
`value`) is the 10M-variable heavy-tailed k-SAT formula: 40M clauses of
length 2 + Geometric(1/2) capped at 30, variable weights rank^(-2/3).  The
round on it is refuted by subsumption in round 0 (the units its
strengthening derives conflict); the reference CPU simplifier stops at the
same place.  `config.stages_run` names the stages that ran.

The JSON line carries:
* roofline: the dominant kernel's algorithmic bytes per launch over its
  average launch time (HIP events on the context's stream), and
  roofline.round: the round's algorithmic bytes B_alg (per stage, for the
  stages that ran; see round_bytes) over the timed step;
* cpu_baseline: the reference CPU simplifier (compiled from its sources into
  oracle/_ref/pfref) on the same full formula, one core per run, three runs
  side by side on three cores, median;
* legs (skip each with its --no-... flag):
  tseitin: the 2M-gate Tseitin circuit (6.8M clauses), same step, with its
    round roofline, the reference's time and output compared with the GPU's;
  heavytail_bve / heavytail_bce_ere: the heavy-tailed formula with SUB off
    (BVE runs and refutes) and with SUB and BVE off (BCE, ERE and GC run), so
    that every stage is timed at 10M variables;
  handoff, ingest, els: result download and device variable map, DIMACS
    ingestion and equivalent-literal substitution on the Tseitin workload;
  batched: 512 formulas resident in HBM, native batch executor against a loop
    of single calls, every output compared;
  serving: BMC sessions under Poisson arrivals through the request server.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload heavytail|tseitin|uniform3]
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


STAGES = ["gc", "ot", "bcp", "lcve", "sort", "sub", "ve", "bce", "ere", "count"]  # pfs_stage_name order
# kernels that only compute the work measures of a profiled step (O_c, D_c,
# O_E, D_E): not part of a round
MEASURE_KERNELS = ("cand_occ_k", "marked_bytes_k", "elected_work_k")


def round_bytes(r, V):
    """Algorithmic bytes of one round, per stage that ran (the round's stage
    mask), from the round's work measures (profiled step).  C, L, V at round
    start, D = 12 C + 4 L (a clause costs a 12-byte header and 4 bytes per
    literal, a list entry or per-variable word 4 bytes); E elected
    variables, O_E / D_E the occurrences and distinct clause bytes of their
    lists, O_c / D_c the same for the election's candidates, R / S the
    resolvent and stack bytes written, D' = 12 C' + 4 L' after the round:
      ot     2D + 24V + 4L   (histogram D + 8V, list offsets 16V, placement D + 4L)
      bcp    D + 12V         (only when units were pending)
      lcve   20V + 4 O_c + D_c
      sort   8 O_E + D_E     (the elected variables' lists: read, keys, write)
      sub    4 O_E + D_E
      ve     2 (4 O_E + D_E) + 56 E + R + S   (decide, slot scans, apply)
      bce    4 O_E + D_E
      ere    4 O_E + D_E     (the stopping round)
      count  D'
      gc     D + D' + 4 C'
    A round that ends early (refuted, or the stopping round) has no C', L':
    D' is taken as D there."""
    C, L = r["clauses_before"], r["literals_before"]
    D = 12 * C + 4 * L
    ran = [STAGES[s] for s in range(len(STAGES)) if (r["stages"] >> s) & 1]
    E, OE, DE = r["elected"], r["elected_occurs"], r["elected_bytes"]
    Oc, Dc = r["cand_occurs"], r["cand_bytes"]
    Ca = r["clauses_after"] or C
    Dp = (12 * r["clauses_after"] + 4 * r["literals_after"]) if r["clauses_after"] else D
    R, S = 4 * r["resolvent_words"], 4 * r["stack_words"]
    model = {"ot": 2 * D + 24 * V + 4 * L, "bcp": D + 12 * V, "lcve": 20 * V + 4 * Oc + Dc, "sort": 8 * OE + DE,
             "sub": 4 * OE + DE, "ve": 2 * (4 * OE + DE) + 56 * E + R + S, "bce": 4 * OE + DE,
             "ere": 4 * OE + DE, "count": Dp, "gc": D + Dp + 4 * Ca}
    return {k: int(model[k]) for k in ran}


def otp_counts(V, C):
    """Range-by-tile counts of the occurrence-table build (otp_plan in
    pfs_kernels.hip): ranges of 4096 literals (8192 past 2048 such ranges),
    tiles of the smallest power of two >= max(4096, 32 x ranges) clauses, at
    most 16,384."""
    n2 = 2 * V + 2
    bits = 12 if (n2 >> 12) < 2048 else 13
    nr = (n2 + (1 << bits) - 1) >> bits
    ct = 4096
    while ct < 32 * nr and ct < 16384:
        ct <<= 1
    return nr * ((C + ct - 1) // ct)


def norm_kernel(name):
    """A kernel's name as the library's profiler reports it: the profilers'
    unsigned template arguments ("mis_step_k<1u>") without the suffix."""
    import re
    name = name.strip("()")
    return re.sub(r"(\d+)u([,>])", r"\1\2", name)


def kernel_bytes(name, rounds, V):
    """Algorithmic bytes per launch of a kernel, from the work measures of
    the profiled step (averaged over the rounds the kernel ran in).

    clause-parallel kernels, per round (C, L, V at round start, D = 12C + 4L):
      hist_k     D + 8V            (read every clause, write the 2V histogram)
      scatter_k  D + 4L + 8V       (read every clause and the list offsets, write one ref per literal)
      count_k    D                 (read every clause)
    occurrence table by literal ranges (N = range-by-tile counts, as the
    library plans them: otp_plan; the tile kernels read a clause's handle,
    its header word and its literals: T = 8C + 4L):
      otp_count_k         T + 4N   (read every clause, write the counts)
      otp_part_k<true>    T + 4N + 8L   (read every clause and the scanned counts, write the pairs)
      otp_part_k<false>   T + 4N + 4L   (the same with literals only)
      otp_range_k<true>   12L + 24V     (read the pairs, write one ref per literal, offsets, sizes, counts)
      otp_range_k<false>  4L + 8V       (read the literals, write the counts)
    election (O_c occurrences in the candidates' lists, D_c their distinct
    clause bytes):
      lcve_classify_k     23V          (histogram 8V, list sizes 8V, states 2V; scores 4V, classes V)
      mis_step_k          4 O_c + D_c  (the first pass: every candidate's lists and clauses)
    per-elected-variable kernels (O_E list entries, D_E clause bytes of the
    elected variables' lists; elected variables share no clause):
      sub_k, ve_decide_k, bce_k   4 O_E + D_E
      ve_apply_k                            4 O_E + D_E + 4 E (decisions) + R + S
    """
    name = norm_kernel(name)
    if name.startswith(("mis_step_k<", "mis_step_dev_k<")):  # any launch shape
        name = name.split("<", 1)[0]
    per = []
    for r in rounds:
        c, l = r["clauses_before"], r["literals_before"]
        d = 12 * c + 4 * l
        tl = 8 * c + 4 * l
        oe, de = r["elected_occurs"], r["elected_bytes"]
        ran = {STAGES[s] for s in range(len(STAGES)) if (r["stages"] >> s) & 1}
        n = otp_counts(V, c)
        if name == "otp_count_k":
            per.append(tl + 4 * n)
        elif name == "otp_part_k<true>":
            per.append(tl + 4 * n + 8 * l)
        elif name == "otp_part_k<false>":
            per.append(tl + 4 * n + 4 * l)
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
        elif name == "lcve_classify_k":
            if "lcve" in ran:
                per.append(23 * V)
        elif name == "mis_step_k":
            if "lcve" in ran and r["cand_occurs"]:
                per.append(4 * r["cand_occurs"] + r["cand_bytes"])
        elif name in ("sub_k", "ve_decide_k", "bce_k"):
            stage = {"sub_k": "sub", "ve_decide_k": "ve", "bce_k": "bce"}[name]
            if oe and stage in ran:
                per.append(4 * oe + de)
        elif name == "ve_apply_k":
            if oe and "ve" in ran:
                per.append(4 * oe + de + 4 * r["elected"] + 4 * r["resolvent_words"] + 4 * r["stack_words"])
        else:
            return None
    if not per or not all(per):
        return None
    return sum(per) / len(per)


def read_counter_csv(path, counter, copy_bytes=256 << 20, reps=3):
    """One rocprofv3 --pmc CSV of this script run with --calibrate-copy:
    ({kernel short name: [bytes per dispatch, ...]}, scale).  The counter
    reads KiB per dispatch; the scale is copy_bytes over what it read for
    the calibration copies, which are the first 'reps' copy dispatches in
    dispatch order (--calibrate-copy runs before anything else copies; the
    workload's own copies later can be larger and must not be taken)."""
    import csv
    import statistics
    vals, copies = {}, []
    with open(path) as f:
        rows = [r for r in csv.DictReader(f) if r["Counter_Name"] == counter]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    for r in rows:
        name = r["Kernel_Name"]
        v = float(r["Counter_Value"]) * 1024.0
        if "copyBuffer" in name or ("copy" in name.lower() and "pfs::" not in name):
            copies.append(v)
            continue
        if not name.startswith(("pfs::", "void pfs::")):
            continue
        short = name.split("pfs::", 1)[1].split("(", 1)[0]
        vals.setdefault(short, []).append(v)
    ref = statistics.median(copies[:reps]) if len(copies) >= reps else 0.0
    return vals, (copy_bytes / ref if ref else None)


def counter_traffic(fetch_csv, write_csv, kernel, copy_bytes=256 << 20):
    """HBM bytes per launch of 'kernel' from two rocprofv3 --pmc runs of this
    script with --calibrate-copy (FETCH_SIZE, then WRITE_SIZE; KiB per
    dispatch), each counter scaled by what it reads for the 256 MiB
    calibration copies.  Returns (bytes per launch, {"fetch_scale": f,
    "write_scale": w})."""
    import statistics
    fv, fs = read_counter_csv(fetch_csv, "FETCH_SIZE", copy_bytes)
    wv, ws = read_counter_csv(write_csv, "WRITE_SIZE", copy_bytes)
    f = [v for k, l in fv.items() if norm_kernel(k) == norm_kernel(kernel) for v in l]
    w = [v for k, l in wv.items() if norm_kernel(k) == norm_kernel(kernel) for v in l]
    if not f or not w or not fs or not ws:
        return None, None
    return statistics.mean(f) * fs + statistics.mean(w) * ws, {"fetch_scale": round(fs, 4), "write_scale": round(ws, 4)}


def write_counter_summary(fetch_csv, write_csv, workload, command, copy_bytes=256 << 20):
    """profiles/<workload>_round_counters.{json,md} from the two rocprofv3
    --pmc runs of this script with --calibrate-copy (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch): per kernel, launches and bytes per launch,
    each counter scaled by what it reads for the 256 MiB calibration copies."""
    import statistics
    fv, fs = read_counter_csv(fetch_csv, "FETCH_SIZE", copy_bytes)
    wv, ws = read_counter_csv(write_csv, "WRITE_SIZE", copy_bytes)
    if not fs or not ws:
        raise SystemExit("bench: no calibration copies in the counter CSVs (run with --calibrate-copy)")
    kernels = {}
    for k in sorted(set(fv) | set(wv)):
        f = statistics.mean(fv[k]) * fs if k in fv else 0.0
        w = statistics.mean(wv[k]) * ws if k in wv else 0.0
        kernels[k] = {"launches": max(len(fv.get(k, [])), len(wv.get(k, []))), "fetch_bytes_per_launch": int(f),
                      "write_bytes_per_launch": int(w), "traffic_bytes_per_launch": int(f + w)}
    out = {"workload": workload, "fetch_scale": round(fs, 4), "write_scale": round(ws, 4), "command": command,
           "kernels": kernels}
    base = os.path.join(ROOT, "profiles", f"{workload}_round_counters")
    with open(base + ".json", "w") as f:
        json.dump(out, f, indent=1)
    rows = sorted(kernels.items(), key=lambda kv: -kv[1]["traffic_bytes_per_launch"])
    with open(base + ".md", "w") as f:
        f.write(f"# HBM traffic per launch: one simplify step on the {workload} workload\n\n")
        f.write("Commands (one MI355X; one rocprofv3 pass per counter; `$REPO` is the repository root):\n\n")
        f.write("```\n" + command + "\n```\n\n")
        f.write(f"Calibration: the 256 MiB device copies of `--calibrate-copy` scale FETCH_SIZE by {fs:.3f} and "
                f"WRITE_SIZE by {ws:.3f}.  Traffic = FETCH_SIZE x {fs:.3f} + WRITE_SIZE x {ws:.3f} per launch.  "
                "bench.py reads the .json beside this file for `roofline.traffic`.\n\n")
        f.write("| kernel | launches | fetch MB/launch | write MB/launch | traffic MB/launch |\n|---|---:|---:|---:|---:|\n")
        for k, v in rows:
            f.write(f"| {k} | {v['launches']} | {v['fetch_bytes_per_launch'] / 1e6:.1f} | "
                    f"{v['write_bytes_per_launch'] / 1e6:.1f} | {v['traffic_bytes_per_launch'] / 1e6:.1f} |\n")
    print(f"bench: wrote {base}.json/.md", file=sys.stderr, flush=True)


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


def host_info():
    """The host's core count and CPU model (lscpu / /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_reference(cnf, extra=(), runs=3, timeout=1500, keep_output=False):
    """The reference CPU simplifier (oracle/_ref/pfref, compiled from the
    reference's src/cpu) on the full formula's DIMACS text: its own parser
    and root propagation, then one simplify call with the step's options.
    'runs' processes side by side, one pinned core each; the median of their
    simplify times.  Returns None when the binary is absent."""
    import statistics
    ref = os.path.join(ROOT, "oracle", "_ref", "pfref")
    if not os.path.exists(ref):
        return None
    ncores = os.cpu_count() or 1
    runs = max(1, min(runs, ncores))
    d = tempfile.mkdtemp(prefix="pfref_")
    try:
        p = os.path.join(d, "w.cnf")
        from parafrost_amd import gen
        gen.write_dimacs(with_units(cnf), p)
        procs = []
        for i in range(runs):
            out = os.path.join(d, f"o{i}.pfd")
            procs.append(subprocess.Popen(["taskset", "-c", str(i), ref, p, os.path.join(d, f"i{i}.pfd"), out]
                                          + REF_OPTIONS + list(extra), stdout=subprocess.PIPE,
                                          stderr=subprocess.DEVNULL, text=True))
        times, parse = [], []
        for pr in procs:
            so, _ = pr.communicate(timeout=timeout)
            line = [x for x in so.splitlines() if x.startswith("{")][-1]
            j = json.loads(line)
            times.append(j["total_s"])
            if j.get("parse_s") is not None:
                parse.append(j["parse_s"])
        from parafrost_amd import pfd
        result = pfd.read(os.path.join(d, "o0.pfd")) if keep_output else None
        status = pfd.read(os.path.join(d, "o0.pfd")).status
    finally:
        import shutil
        shutil.rmtree(d, ignore_errors=True)
    med = statistics.median(times)
    return {"seconds": med, "runs_s": [round(t, 3) for t in times], "parse_seconds":
            statistics.median(parse) if parse else None, "status": ["UNSAT", "SAT", "UNSOLVED"][status],
            "result": result, "runs": runs}


def cpu_baseline(cnf, name):
    """cpu_baseline of the JSON line: the reference on the full workload
    (three runs side by side, one core each, median), or the oracle
    restatement on one core when the reference binary is absent."""
    info = host_info()
    r = cpu_reference(cnf)
    if r is not None:
        return {"value": cnf.nclauses / r["seconds"], "unit": "clauses/s", "cores": 1, "kind": "reference",
                "sample": (f"full {name} workload ({cnf.nclauses} clauses); reference CPU simplifier (ParaFROST "
                           f"src/cpu, compiled from its sources) simplify time, median of {r['runs']} runs side by "
                           f"side, one pinned core each; the reference's own parse excluded"),
                "seconds": round(r["seconds"], 3), "runs_s": r["runs_s"], "status": r["status"],
                "parse_seconds": r["parse_seconds"], **info}
    from oracle import oracle
    t = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=0, phases=1,
                        bce_en=1)
    return {"value": cnf.nclauses / t.seconds, "unit": "clauses/s", "cores": 1, "kind": "port",
            "sample": f"full {name} workload; oracle restatement simplify time {t.seconds:.3f} s, one core",
            "seconds": round(t.seconds, 3), **info}


def step_loop(s, steps, warmup):
    """warm-up calls, then 'steps' timed simplify calls on the resident input
    (device copy restored inside the timed region)."""
    status = None
    for _ in range(warmup):
        s.rewind()
        status = s.simplify()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.rewind()
        status = s.simplify()
    sync()
    return status, (time.perf_counter() - t0) / steps


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def committed_traffic(workload, kernel):
    """HBM bytes per launch of 'kernel' on 'workload' from the committed
    counter summary profiles/<workload>_round_counters.json (written from two
    rocprofv3 --pmc runs of this script, FETCH_SIZE and WRITE_SIZE, each
    scaled by what it reads for a device copy of known size; the .md beside
    it gives the commands).  (None, None) when there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_round_counters.json")
    try:
        with open(path) as f:
            j = json.load(f)
    except OSError:
        return None, None
    k = {norm_kernel(n): v for n, v in j.get("kernels", {}).items()}.get(norm_kernel(kernel))
    if not k:
        return None, None
    return k["traffic_bytes_per_launch"], {"fetch_scale": j.get("fetch_scale"), "write_scale": j.get("write_scale"),
                                          "source": os.path.relpath(path, ROOT)}


def profile_step(s, cnf, dt, traffic_csv=None, workload=None):
    """One event-profiled step: per-kernel times, the rounds' work measures,
    the round roofline and the dominant kernel's roofline."""
    s.set_option("profilegpu", 1)
    s.rewind()
    st = s.simplify()
    kt = {norm_kernel(k): v for k, v in s.kernel_times().items() if k not in MEASURE_KERNELS}
    stages = s.stage_times()
    s.set_option("profilegpu", 0)
    rounds = s.result(st).rounds
    V = cnf.maxvar
    per_stage = {}
    for r in rounds:
        for k, v in round_bytes(r, V).items():
            per_stage[k] = per_stage.get(k, 0) + v
    b_alg = sum(per_stage.values())
    stages_run = [k for k in STAGES if k in per_stage]
    dom = max(kt.items(), key=lambda kv: kv[1][0], default=None)
    modelled = {k: v for k, v in kt.items() if kernel_bytes(k, rounds, V)}
    roof_kernel = dom if dom and dom[0] in modelled else max(modelled.items(), key=lambda kv: kv[1][0], default=None)
    roof = {"bound": "hbm"}
    if roof_kernel:
        name, (ms, n) = roof_kernel
        avg_s = ms / n / 1e3
        runs_in = sum(1 for r in rounds if kernel_bytes(name, [r], V))
        b = kernel_bytes(name, rounds, V) * runs_in / n  # bytes of the rounds it ran in, spread over its launches
        ach = b / avg_s / 1e9
        traffic, scales = None, None
        if traffic_csv:
            fcsv, wcsv = traffic_csv.split(",")
            traffic, scales = counter_traffic(fcsv, wcsv, name)
        elif workload:
            traffic, scales = committed_traffic(workload, name)
        roof.update({"kernel": name, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic else None,
                     "traffic_scales": scales, "avg_launch_us": round(avg_s * 1e6, 2),
                     "algorithmic_bytes_per_launch": int(b)})
    roof["dominant_kernel"] = dom[0] if dom else None
    roof["dominant_kernel_ms_per_step"] = round(dom[1][0], 3) if dom else None
    ach_round = b_alg / dt / 1e9
    # every kernel of the profiled step (the work-measure kernels left out):
    # what the timed step spends beyond it is host reads of counts,
    # launch-to-launch latency and the rewind's device copies.  The profiled
    # step records events around every launch, which stretches its kernels
    # a little, so on a round with almost no time outside kernels the
    # difference can come out slightly negative.
    kernel_ms = sum(v[0] for v in kt.values())
    roof["round"] = {"achieved": round(ach_round, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_round / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(b_alg),
                     "ms_per_step": round(dt * 1e3, 3), "kernel_ms_per_step": round(kernel_ms, 3),
                     "outside_kernels_ms": round(dt * 1e3 - kernel_ms, 3), "bytes_per_stage": per_stage}
    mk = {}
    for k, v in sorted(modelled.items(), key=lambda kv: -kv[1][0]):
        runs_in = sum(1 for r in rounds if kernel_bytes(k, [r], V))
        b = kernel_bytes(k, rounds, V) * runs_in / v[1]
        mk[k] = {"frac": round(b / (v[0] / v[1] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                 "avg_launch_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1]}
    roof["modelled_kernels"] = mk
    print("bench: kernels " + json.dumps({k: [round(v[0], 3), v[1]] for k, v in
                                          sorted(kt.items(), key=lambda kv: -kv[1][0])[:16]}),
          file=sys.stderr, flush=True)
    print("bench: stages " + json.dumps({k: [round(v[0], 3), v[1]] for k, v in stages.items()}),
          file=sys.stderr, flush=True)
    return roof, stages_run, rounds, kt


def leg(cnf, name, opts, steps=3, warmup=1, extra_ref=(), ref=False, compare_ref=False):
    """A single-formula leg: ms per step with inputs resident, its round
    roofline, the stages that ran and (with ref) the reference's time on the
    same formula and options; with compare_ref the GPU result is compared
    with the reference's output (status, clauses, units as a set, stack,
    variable states)."""
    import numpy as np
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0, keep_input=True, **opts)
    try:
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
        status, dt = step_loop(s, steps, warmup)
        shape = s.launch_shape()
        roof, stages_run, rounds, _ = profile_step(s, cnf, dt)
        res = s.result(status)
    except Exception:
        s.close()
        raise
    out = {"workload": name, "variables": cnf.maxvar, "clauses": cnf.nclauses, "literals": cnf.nlits,
           "options": " ".join(f"{k}={v}" for k, v in opts.items()), "ms_per_step": round(dt * 1e3, 3),
           "clauses_per_s": round(cnf.nclauses / dt, 1), "round_result": ["UNSAT", "SAT", "UNSOLVED"][status],
           "rounds": len(rounds), "elected": [x["elected"] for x in rounds],
           "sub_passes": [x["sub_passes"] for x in rounds], "stages_run": stages_run,
           "round_roofline": roof["round"], "dominant_kernel": roof.get("dominant_kernel"), "launch_shape": shape}
    if ref:
        r = cpu_reference(cnf, extra_ref, runs=1, keep_output=compare_ref)
        if r is not None:
            out["cpu_reference_s"] = round(r["seconds"], 3)
            out["cpu_reference_parse_s"] = r["parse_seconds"]
            out["cpu_reference_status"] = r["status"]
            out["speedup_vs_reference"] = round(r["seconds"] / dt, 1)
            if compare_ref and r["result"] is not None:
                o = r["result"]
                same = res.status == o.status
                if same and o.status != 0:
                    same = (np.array_equal(res.sizes, o.sizes) and np.array_equal(res.lits, o.lits)
                            and np.array_equal(res.flags, o.flags) and np.array_equal(res.resolved, o.resolved)
                            and np.array_equal(res.vstate, o.vstate)
                            and np.array_equal(np.sort(res.units), np.sort(o.units)))
                out["identical_to_reference"] = bool(same)
    print("bench: leg " + json.dumps(out), file=sys.stderr, flush=True)
    return out, s


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


def bench_batched(n=512, workers=16):
    """Batched workload: n formulas (clause counts log-uniform in 1k-500k,
    Tseitin or heavy-tailed) resident in HBM, simplified once as a loop of
    plain single-formula calls and once by pfs_batch_simplify with 'workers'
    native threads (one stream per formula).  formulas/s of both, per-formula
    latency percentiles of the batched run, and a bit-exact comparison of
    every output (and proof-free result arrays) with the loop's."""
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
    ref = [ctx[i].result(loop_res[i]) for i in range(n)]
    for s in ctx:
        s.rewind()
    t = time.perf_counter()
    res, ms = batch_simplify(ctx, workers)
    batch_s = time.perf_counter() - t
    same = 0
    for i in range(n):
        b = ctx[i].result(res[i])
        same += int(b.status == ref[i].status and np.array_equal(b.lits, ref[i].lits) and
                    np.array_equal(b.sizes, ref[i].sizes) and np.array_equal(b.resolved, ref[i].resolved) and
                    np.array_equal(b.flags, ref[i].flags) and np.array_equal(b.vstate, ref[i].vstate) and
                    np.array_equal(b.units, ref[i].units))
    for s in ctx:
        s.close()
    out = {"formulas": n, "clauses": ncl, "workers": workers,
           "formulas_per_s": round(n / batch_s, 1), "loop_formulas_per_s": round(n / loop_s, 1),
           "speedup_vs_loop": round(loop_s / batch_s, 2), "clauses_per_s": round(ncl / batch_s, 1),
           "p50_ms": round(float(np.percentile(ms, 50)), 3), "p99_ms": round(float(np.percentile(ms, 99)), 3),
           "batch_s": round(batch_s, 3), "loop_s": round(loop_s, 3), "identical": f"{same}/{n}"}
    print("bench: batched " + json.dumps(out), file=sys.stderr, flush=True)
    return out


def bench_serving(rates=(10, 20, 50, 100, 200, 300, 500), nsessions=16, workers=8, requests=1000,
                  low_rate_requests=200):
    """Serving workload: 16 BMC sessions (64 latches, 2,000 gates, 10 frames
    per request, 1-32 assumptions) whose formulas stay resident.  Every
    session is first advanced to a depth drawn uniformly from 10-190 frames
    (untimed), so that the measured requests act on formulas of every depth
    the workload names; then requests (each session's next ones, up to its
    200-frame bound, after which it starts over from a fresh depth) arrive
    as a Poisson stream at each rate and are served by the request server
    (continuous batching over 'workers' native threads).  Latency =
    completion - arrival; 'requests' per rate (fewer at the two lowest
    rates, which would otherwise take minutes)."""
    import numpy as np
    from parafrost_amd import bmc
    from parafrost_amd.simplifier import Server, Simplifier
    t0 = time.time()
    rng0 = np.random.default_rng(0x5EED0005)
    sweep = []
    for rate in rates:
        nreq = low_rate_requests if rate < 50 else requests
        rng = np.random.default_rng(rate)
        # sessions: a circuit each, pre-advanced to a random depth; a session
        # that reaches 200 frames starts over as a fresh session (a new
        # context, kept until the rate's requests are done)
        ctx, unroll = [], []
        cur = list(range(nsessions))

        def fresh():
            s = Simplifier(0, **OPTIONS)
            s.upload(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
            ctx.append(s)
            unroll.append(bmc.Unrolling(bmc.Circuit(int(rng0.integers(1 << 30)))))
            return len(ctx) - 1

        for i in range(nsessions):
            k = fresh()
            d0 = int(rng.integers(0, 20)) * 10
            while unroll[k].frames < d0:
                rq = unroll[k].request(10, rng)
                ctx[k].append(rq.maxvar, rq.offsets, rq.lits)
                ctx[k].assume(rq.assumed)
                ctx[k].simplify()
        # the measured requests, prepared ahead of the arrivals
        order = []
        for _ in range(nreq):
            slot = int(rng.integers(nsessions))
            if unroll[cur[slot]].frames >= 200:
                cur[slot] = fresh()
            k = cur[slot]
            order.append((k, unroll[k].request(10, rng), unroll[k].frames))
        times = bmc.poisson_schedule(len(order), rate, rate + 1)
        srv = Server(workers)
        tickets = []
        for (k, rq, _), rel in zip(order, times):
            tickets.append(srv.submit(ctx[k], float(rel), rq.maxvar, rq.offsets, rq.lits, rq.assumed))
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
        frames = np.array([f for _, _, f in order])
        row = {"rate_rps": rate, "requests": len(order), "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(np.max(lat)), 3),
               "throughput_rps": round(len(order) / (span / 1e3), 1),
               "frames_min_median_max": [int(frames.min()), int(np.median(frames)), int(frames.max())],
               "contexts": len(ctx)}
        sweep.append(row)
        print("bench: serving " + json.dumps(row), file=sys.stderr, flush=True)
    print(f"bench: serving done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    return {"sessions": nsessions, "workers": workers, "frames_per_request": 10, "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="heavytail", choices=["heavytail", "tseitin", "uniform3"])
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="extra simplifier option (launch shaping such as workorder=0)")
    ap.add_argument("--traffic-csv", default=None, metavar="FETCH.csv,WRITE.csv",
                    help="rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE counter files of this workload "
                         "(runs with --calibrate-copy) for roofline.traffic")
    ap.add_argument("--no-tseitin", action="store_true", help="skip the Tseitin leg (and handoff, ingest)")
    ap.add_argument("--no-stage-legs", action="store_true", help="skip the heavy-tailed BVE and BCE/ERE legs")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-ingest", action="store_true", help="skip the DIMACS ingestion leg")
    ap.add_argument("--no-els", action="store_true", help="skip the equivalent-literal substitution leg")
    ap.add_argument("--no-legs", action="store_true", help="skip every leg (headline only)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run 256 MiB device copies first (reference kernel for counter runs)")
    ap.add_argument("--write-counter-summary", default=None, metavar="FETCH.csv,WRITE.csv",
                    help="write profiles/<workload>_round_counters.{json,md} from two counter runs and exit")
    ap.add_argument("--counter-command", default="", help="the commands recorded in that summary")
    args = ap.parse_args()
    if args.write_counter_summary:
        fcsv, wcsv = args.write_counter_summary.split(",")
        write_counter_summary(fcsv, wcsv, args.workload, args.counter_command)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(formulas shard by independent instances or elected variables; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.no_legs:
        args.no_tseitin = args.no_stage_legs = args.no_batched = args.no_serving = True
        args.no_ingest = args.no_els = True

    from parafrost_amd import gen
    from parafrost_amd.simplifier import Simplifier

    t0 = time.time()
    cnf = gen.workload(args.workload, scale=args.scale)
    print(f"bench: workload {args.workload} scale {args.scale}: {cnf.maxvar} vars, {cnf.nclauses} clauses, "
          f"{cnf.nlits} literals ({time.time() - t0:.1f}s to generate)", file=sys.stderr, flush=True)
    if args.calibrate_copy:
        calibrate_copy()
    opts = dict(OPTIONS)
    for kv in args.set:
        k, v = kv.split("=", 1)
        opts[k] = int(v)
    s = Simplifier(0, keep_input=True, **opts)
    s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
    status, dt = step_loop(s, args.steps, args.warmup)
    res = s.result(status)
    shape = s.launch_shape()
    print(f"bench: {args.steps} steps, {dt * 1e3:.2f} ms/step, result {res.status}, "
          f"{res.sizes.shape[0]} clauses out, rounds {res.rounds}", file=sys.stderr, flush=True)
    roofline, stages_run, rounds, _ = profile_step(s, cnf, dt, args.traffic_csv,
                                                    args.workload if args.scale == 1.0 else None)
    s.close()

    legs = {}
    tcnf = None
    if not args.no_tseitin:
        tcnf = cnf if args.workload == "tseitin" and args.scale == 1.0 else gen.workload("tseitin")
        legs["tseitin"], ts = leg(tcnf, "tseitin", OPTIONS, ref=not args.no_cpu_baseline, compare_ref=True)
        try:
            legs["handoff"] = bench_handoff(ts, tcnf)
        finally:
            ts.close()
        if not args.no_ingest:
            legs["ingest"] = bench_ingest(tcnf)
    if not args.no_els:
        legs["els"] = bench_els()
    if not args.no_stage_legs and args.workload == "heavytail":
        hv = cnf
        for key, o, extra in (("heavytail_bve", {"sub": 0, "ve+": 0}, ["-no-sub", "-no-ve+"]),
                              ("heavytail_bce_ere", {"ve": 0, "sub": 0, "ve+": 0}, ["-no-ve", "-no-sub", "-no-ve+"])):
            legs[key], ls = leg(hv, key, dict(OPTIONS, **o), ref=False)
            ls.close()
    if not args.no_batched:
        legs["batched"] = bench_batched()
    if not args.no_serving:
        legs["serving"] = bench_serving()

    base = None
    if not args.no_cpu_baseline:
        try:
            base = cpu_baseline(cnf, args.workload)
        except Exception as e:  # the baseline is reported, never required
            print(f"bench: cpu baseline failed: {e}", file=sys.stderr)
    if base and "status" in base:
        base["same_status_as_gpu"] = base["status"] == ["UNSAT", "SAT", "UNSOLVED"][status]
    if "ingest" in legs and legs.get("tseitin", {}).get("cpu_reference_parse_s") is not None:
        legs["ingest"]["cpu_reference_parse_s"] = legs["tseitin"]["cpu_reference_parse_s"]
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
                   "clauses": cnf.nclauses, "literals": cnf.nlits,
                   "options": " ".join(f"{k}={v}" for k, v in opts.items()) + " (SUB+BVE+BCE, ERE)",
                   "round_result": ["UNSAT", "SAT", "UNSOLVED"][status], "rounds": len(rounds),
                   "stages_run": stages_run, "elected": [x["elected"] for x in rounds],
                   "sub_passes": [x["sub_passes"] for x in rounds], "launch_shape": shape,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": base,
    }
    out.update(legs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
