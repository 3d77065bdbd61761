"""Writes tests/golden/fullsize.json: digests (tests.common.result_digest) of
the oracle's deterministic-mode outputs on the full-size workloads -- the
2M-gate Tseitin circuit and the 10M-variable heavy-tailed formula -- under
the option sets the benchmark times, with proofs.  The GPU tests compare
their outputs with these digests (tests/test_gpu_fullsize.py).

Each entry also records the reference CPU simplifier (oracle/_ref/pfref,
built from /root/reference by oracle/build_ref.sh) on the same formula: its
input state must equal the generator's prepared formula, and its output
must equal the oracle's in REF mode (and is compared with DET mode).
Takes about 20 minutes on one core per workload leg set.

Usage: python tests/golden/make_fullsize.py [tseitin|heavytail ...]
"""
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from parafrost_amd import gen, pfd  # noqa: E402
from tests.common import result_digest  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "pfref")
OUT = os.path.join(HERE, "fullsize.json")

# (name, workload, simplifier options (binding names), reference options)
LEGS = [
    ("tseitin.p1bce", "tseitin", {"phases": 1, "bce": 1}, ["--phases=1", "-bce"]),
    ("tseitin.default", "tseitin", {}, []),
    ("heavytail.p1bce", "heavytail", {"phases": 1, "bce": 1}, ["--phases=1", "-bce"]),
    ("heavytail.p1bce_nosub", "heavytail", {"phases": 1, "bce": 1, "sub": 0, "ve+": 0},
     ["--phases=1", "-bce", "-no-sub", "-no-ve+"]),
    ("heavytail.p1bce_bce_ere", "heavytail", {"phases": 1, "bce": 1, "sub": 0, "ve+": 0, "ve": 0},
     ["--phases=1", "-bce", "-no-sub", "-no-ve+", "-no-ve"]),
]


def oracle_opts(o):
    from tests.common import oracle_kw
    return oracle_kw(o)


def with_units(cnf):
    units = np.asarray(cnf.trail, np.uint32)
    if not units.shape[0]:
        return cnf
    lits = np.concatenate([cnf.lits, units])
    off = np.concatenate([cnf.offsets, cnf.offsets[-1] + 1 + np.arange(units.shape[0], dtype=np.uint64)])
    return gen.Cnf(cnf.maxvar, off, lits)


def run_reference(cnf, ropts, d):
    p = os.path.join(d, "w.cnf")
    if not os.path.exists(p):
        gen.write_dimacs(with_units(cnf), p)
    drat = os.path.join(d, "p.drat")
    t = time.time()
    r = subprocess.run([REF, p, os.path.join(d, "i.pfd"), os.path.join(d, "o.pfd"), "-proof",
                        f"--proofout={drat}"] + ropts, capture_output=True, text=True, timeout=3600)
    secs = time.time() - t
    line = [x for x in r.stdout.splitlines() if x.startswith("{")]
    out = pfd.read(os.path.join(d, "o.pfd"))
    with open(os.path.join(d, "o.pfd.proofpos")) as f:
        pos = int(f.read())
    with open(drat, "rb") as f:
        out.proof = f.read()[pos:]
    os.remove(drat)
    return pfd.read(os.path.join(d, "i.pfd")), out, (json.loads(line[-1]) if line else {}), secs


def main():
    want = set(sys.argv[1:]) or {"tseitin", "heavytail"}
    entries = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            entries = {e["name"]: e for e in json.load(f)}
    cache = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, wl, opts, ropts in LEGS:
            if wl not in want:
                continue
            if wl not in cache:
                cache.clear()
                t = time.time()
                cache[wl] = gen.workload(wl)
                print(f"{wl}: generated in {time.time() - t:.1f}s", flush=True)
            cnf = cache[wl]
            vs = np.zeros(cnf.maxvar + 1, np.uint8)
            vs[np.asarray(cnf.trail, np.int64) >> 1] = 2
            t = time.time()
            det = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1,
                                  vstate=vs, proof_en=1, **oracle_opts(opts))
            det_s = time.time() - t
            e = {"name": name, "workload": wl, "variables": cnf.maxvar, "clauses": cnf.nclauses,
                 "literals": cnf.nlits, "options": opts, "oracle_det": result_digest(det),
                 "oracle_det_seconds": round(det_s, 1)}
            print(name, "det", e["oracle_det"]["status"], f"{det_s:.1f}s", flush=True)
            if os.path.exists(REF):
                d = os.path.join(tmp, wl)
                os.makedirs(d, exist_ok=True)
                inp, out, rj, rs = run_reference(cnf, ropts, d)
                ref_mode = oracle.simplify_state(inp, det=0, proof_en=1, **oracle_opts(opts))
                prepared_same = (np.array_equal(np.sort(inp.units), np.sort(np.asarray(cnf.trail, np.uint32)))
                                 and inp.sizes.shape[0] == cnf.nclauses and inp.lits.shape[0] == cnf.nlits)
                rd, od = result_digest(out), result_digest(ref_mode)
                e["reference"] = {"digest": rd, "seconds": round(rs, 1), "simplify_seconds": rj.get("total_s"),
                                  "input_is_prepared_formula": bool(prepared_same),
                                  "oracle_ref_equal": rd == od, "oracle_det_equal": rd == e["oracle_det"]}
                print(name, "reference", rd["status"], "ref==oracle(REF):", rd == od, "ref==oracle(DET):",
                      rd == e["oracle_det"], flush=True)
            entries[name] = e
            with open(OUT, "w") as f:
                json.dump([entries[n] for n, *_ in LEGS if n in entries], f, indent=1)


if __name__ == "__main__":
    main()
