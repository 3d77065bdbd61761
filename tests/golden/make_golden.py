"""Writes the golden fixtures of this directory with the reference CPU
simplifier (ParaFROST src/cpu, compiled from its own sources by
oracle/build_ref.sh into oracle/_ref/pfref).

Each fixture is a pair of PFD1 dumps (parafrost_amd/pfd.py) and a proof:
  <case>.in.pfd.gz        -- the simplifier input the reference extracted (after
                          its parser, root BCP and rootify/shrinkTop)
  <case>.<opts>.out.pfd.gz -- the reference's output: clauses in simplifier order,
                    new root units, eliminated-variable stack, variable states
  <case>.<opts>.proof.gz  -- the binary DRAT lines the reference simplifier
                    wrote (run with -proof; the part after rootify/shrinkTop)
  <case>.<opts>.map.gz    -- the state after the reference's host write-back
                    (Solver::map(true) when canMap(), else newBeginning();
                    PFM1, oracle/vmap.py read_pfm), for unsolved runs
manifest.json lists every fixture with the reference options it was run
with; dimacs_manifest.json the parser fixtures (main_dimacs).  Run from the repository root:  python tests/golden/make_golden.py
"""
import gzip
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from parafrost_amd import gen  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "pfref")

# reference option strings (src/cpu/options.cpp) -> keyword names used by the
# oracle wrapper and the GPU binding
OPTSETS = {
    "default": ([], {}),
    "p1bce": (["--phases=1", "-bce"], {"phases": 1, "bce": 1}),
    "all": (["-all"], {"all": 1}),
    "all_mu10": (["-all", "--mupos=10", "--muneg=10", "--phases=3"], {"all": 1, "mupos": 10, "muneg": 10, "phases": 3}),
}


def cases():
    """(name, formula, learnt): learnt None, or "<seed>,<percent>" for the
    reference harness's PFREF_LEARNT (that share of the parsed clauses handed
    over as learnt clauses, with lbd and usage)."""
    yield "u3_n60", gen.uniform_ksat(60, 256, 3, seed=11), None
    yield "u3_n300", gen.uniform_ksat(300, 1278, 3, seed=12), None
    yield "u3_n2000", gen.uniform_ksat(2000, 8520, 3, seed=13), None
    yield "u3lo_n1000", gen.uniform_ksat(1000, 3000, 3, seed=14), None
    yield "u2_n400", gen.uniform_ksat(400, 500, 2, seed=15), None
    yield "ts_g300", gen.tseitin(300, 20, 3, seed=21), None
    yield "ts_g2000", gen.tseitin(2000, 60, 10, seed=22), None
    yield "ts_g4000", gen.tseitin(4000, 100, 12, seed=23), None
    yield "ht_v300", gen.heavytail(300, 1200, seed=31), None
    yield "ht_v2000", gen.heavytail(2000, 8000, seed=32), None
    yield "ht_v3000", gen.heavytail(3000, 12000, seed=33), None
    yield "lt_u3_n2000", gen.uniform_ksat(2000, 8520, 3, seed=13), "41,30"
    yield "lt_ts_g4000", gen.tseitin(4000, 100, 12, seed=23), "42,30"
    yield "lt_ht_v3000", gen.heavytail(3000, 12000, seed=33), "43,40"


def messy_text(seed, nvars, nclauses, k=3, units=4, unsat=None, percent=False, crlf=False):
    """A DIMACS text that exercises the parser: comment lines before the
    header and between clauses, comments after a clause's 0, blank lines,
    tabs and several blanks, CRLF line ends, '+' signs, leading zeros,
    clauses spread over lines, duplicate literals, tautologies, unit clauses
    (some repeated, some satisfying or shrinking later clauses) and an
    optional SATLIB '%' ending.  unsat: None, "empty" (an empty clause),
    "units" (opposite units) or "bcp" (units that propagate to a conflict)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cls = []
    for _ in range(nclauses):
        vs = rng.choice(np.arange(1, nvars + 1), size=k, replace=False)
        cl = [int(v) if rng.random() < 0.5 else -int(v) for v in vs]
        r = rng.random()
        if r < 0.08:
            cl.insert(int(rng.integers(0, len(cl) + 1)), cl[int(rng.integers(0, len(cl)))])  # duplicate
        elif r < 0.12:
            cl.insert(int(rng.integers(0, len(cl) + 1)), -cl[int(rng.integers(0, len(cl)))])  # tautology
        cls.append(cl)
    for _ in range(units):
        u = int(rng.integers(1, nvars + 1)) * (1 if rng.random() < 0.5 else -1)
        at = int(rng.integers(0, len(cls) + 1))
        cls.insert(at, [u])
        if rng.random() < 0.5:
            cls.insert(int(rng.integers(at + 1, len(cls) + 1)), [u, u])  # the same unit again
    if unsat == "empty":
        cls.insert(int(rng.integers(0, len(cls) + 1)), [])
    elif unsat == "units":
        u = int(rng.integers(1, nvars + 1))
        cls.insert(int(rng.integers(0, len(cls) // 2)), [u])
        cls.insert(int(rng.integers(len(cls) // 2, len(cls) + 1)), [-u])
    elif unsat == "bcp":
        a, b = (int(x) for x in rng.choice(np.arange(1, nvars + 1), 2, replace=False))
        cls += [[-a, b], [-a, -b], [a]]
    nl = "\r\n" if crlf else "\n"
    out = ["c messy text, seed %d%s" % (seed, nl), "c p cnf 1 1 is only a comment%s" % nl, nl,
           "p cnf  %d\t%d %s" % (nvars, len(cls), nl)]

    def num(x):
        t = str(abs(x))
        if rng.random() < 0.05:
            t = "0" + t
        if x < 0:
            return "-" + t
        return ("+" + t) if rng.random() < 0.1 else t

    for cl in cls:
        if rng.random() < 0.04:
            out.append("c a comment line between clauses 12 -3 0 %% x%s" % nl)
        if rng.random() < 0.03:
            out.append(nl)
        parts = [num(x) for x in cl] + ["0"]
        line = ""
        for j, t in enumerate(parts):
            line += t
            if j + 1 < len(parts):
                line += nl if rng.random() < 0.05 else (" \t"[int(rng.integers(0, 2))] * int(rng.integers(1, 3)))
        if rng.random() < 0.03:
            line += " c trailing comment"
        out.append(line + nl)
    if percent:
        out.append("%" + nl + "0" + nl)
    return "".join(out).encode()


def dimacs_cases():
    """(name, text) of the parser fixtures; the last ones are rejected."""
    yield "mx_u3_a", messy_text(101, 60, 200)
    yield "mx_u3_b", messy_text(102, 300, 1200, units=12, crlf=True)
    yield "mx_u3_c", messy_text(103, 2000, 8000, units=40, percent=True)
    yield "mx_u4_d", messy_text(104, 500, 1500, k=4, units=25)
    yield "mx_u2_e", messy_text(105, 400, 450, k=2, units=4)
    yield "mx_unsat_empty", messy_text(106, 100, 300, unsat="empty")
    yield "mx_unsat_units", messy_text(107, 100, 300, unsat="units")
    yield "mx_unsat_bcp", messy_text(108, 100, 300, unsat="bcp")
    yield "err_var", b"p cnf 3 1\n1 4 0\n"
    yield "err_token", b"p cnf 3 1\n1 x 0\n"
    yield "err_noheader", b"1 2 0\n"
    yield "err_unterminated", b"p cnf 3 1\n1 2"
    yield "err_clauses", b"p cnf 3 1\n1 2 0\n2 3 0\n"
    yield "err_zerovars", b"p cnf 0 1\n"


def main_dimacs():
    """Parser fixtures: <name>.dimacs (the text), <name>.in.pfd.gz (the
    reference's simplifier input after its parse, BCP and shrinkTop),
    <name>.p1bce.out.pfd.gz (its simplifier output with --phases=1 -bce),
    listed in dimacs_manifest.json with the outcome: "ok", "unsat" (the text
    refutes itself) or "error" (the reference rejects it)."""
    manifest = []
    for name, text in dimacs_cases():
        path = os.path.join(HERE, name + ".dimacs")
        with open(path, "wb") as f:
            f.write(text)
        inp = os.path.join(HERE, name + ".in.pfd")
        out = os.path.join(HERE, name + ".p1bce.out.pfd")
        for x in (inp, out):
            if os.path.exists(x):
                os.remove(x)
        r = subprocess.run([REF, path, inp, out, "--phases=1", "-bce"], capture_output=True, text=True, timeout=600)
        entry = {"name": name, "dimacs": name + ".dimacs", "options": {"phases": 1, "bce": 1}}
        if r.returncode != 0:
            entry["outcome"] = "error"
            msg = [ln for ln in (r.stdout + r.stderr).splitlines() if "ERROR" in ln]
            entry["reference_error"] = msg[0].split("ERROR - ", 1)[-1].strip() if msg else ""
        elif not os.path.exists(inp):
            entry["outcome"] = "unsat"
        else:
            entry["outcome"] = "ok"
            gz(inp)
            entry["input"] = name + ".in.pfd.gz"
        if os.path.exists(out) and entry["outcome"] == "ok":
            gz(out)
            entry["output"] = name + ".p1bce.out.pfd.gz"
        elif os.path.exists(out):
            os.remove(out)
        manifest.append(entry)
        print(name, entry["outcome"], entry.get("reference_error", ""))
    with open(os.path.join(HERE, "dimacs_manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


def els_cases():
    """(name, formula, rounds) of the ELS fixtures; rounds None: the
    reference's default (decomposemin)."""
    yield "els_eq_a", gen.equivalences(200, 500, 10, 4, seed=1), None
    yield "els_eq_b", gen.equivalences(2000, 6000, 80, 6, seed=2, units=5), None
    yield "els_eq_anti", gen.equivalences(1000, 3000, 60, 5, seed=16, anti=6), None
    yield "els_eq_anti_r1", gen.equivalences(1000, 3000, 60, 5, seed=16, anti=6), 1
    yield "els_eq_anti_r5", gen.equivalences(1000, 3000, 60, 5, seed=16, anti=6), 5
    yield "els_eq_anti_unsat", gen.equivalences(1000, 3000, 60, 5, seed=6, anti=40), None
    yield "els_eq_units", gen.equivalences(1500, 4000, 100, 4, seed=7, anti=10, units=3), None
    yield "els_eq_units_b", gen.equivalences(1500, 4000, 100, 4, seed=17, anti=10, units=2), None
    yield "els_eq_conflict", gen.equivalences(500, 1500, 30, 5, seed=3, conflict=True), None
    yield "els_eq_2sat", gen.equivalences(800, 300, 40, 5, k=2, seed=8, anti=10), None
    yield "els_eq_big", gen.equivalences(20000, 60000, 1500, 5, seed=9, anti=200, units=20), None
    yield "els_u3_n2000", gen.uniform_ksat(2000, 8520, 3, seed=13), None
    yield "els_ts_g2000", gen.tseitin(2000, 60, 10, seed=22), None
    yield "els_u2_n400", gen.uniform_ksat(400, 500, 2, seed=15), None
    yield "els_ht_v2000", gen.heavytail(2000, 8000, seed=32), None


def main_els():
    """ELS fixtures: <name>.in.pfd.gz (the reference's root-clean input),
    <name>.els.pfd.gz (its state after decompose rounds + rootify +
    shrinkTop, PFREF_ELS_OUT of oracle/ref_harness.cpp), listed in
    els_manifest.json."""
    manifest = []
    for name, cnf, rounds in els_cases():
        cnf_path = os.path.join(HERE, name + ".cnf")
        gen.write_dimacs(cnf, cnf_path)
        inp = os.path.join(HERE, name + ".in.pfd")
        out = os.path.join(HERE, name + ".els.pfd")
        env = dict(os.environ, PFREF_ELS_OUT=out)
        if rounds:
            env["PFREF_ELS_ROUNDS"] = str(rounds)
        with tempfile.TemporaryDirectory() as side:  # the simplify output of the run is not a fixture
            r = subprocess.run([REF, cnf_path, inp, os.path.join(side, name + ".out.pfd")], capture_output=True,
                               text=True, timeout=600, env=env)
        os.remove(cnf_path)
        entry = {"name": name, "rounds": rounds}
        if not os.path.exists(inp):
            print("skip (refuted at parse time)", name)
            continue
        gz(inp)
        gz(out)
        entry.update({"input": name + ".in.pfd.gz", "output": name + ".els.pfd.gz"})
        line = [x for x in r.stdout.splitlines() if x.startswith("{")]
        entry["reference"] = json.loads(line[-1]) if line else None
        if entry["reference"]:
            entry["reference"].pop("els_s", None)  # a timing, not part of the fixture
        manifest.append(entry)
        print(name, entry["reference"])
    with open(os.path.join(HERE, "els_manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


def gz(path):
    """Compresses path to path.gz (no timestamp in the header) unless an
    existing path.gz already holds the same bytes."""
    with open(path, "rb") as a:
        data = a.read()
    os.remove(path)
    if os.path.exists(path + ".gz"):
        with gzip.open(path + ".gz", "rb") as old:
            if old.read() == data:
                return
    with open(path + ".gz", "wb") as f, gzip.GzipFile(filename="", mode="wb", fileobj=f, compresslevel=9,
                                                       mtime=0) as b:
        b.write(data)


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/pfref is not built (run oracle/build_ref.sh)")
    manifest = []
    for name, cnf, learnt in cases():
        cnf_path = os.path.join(HERE, name + ".cnf")
        gen.write_dimacs(cnf, cnf_path)
        inp = os.path.join(HERE, name + ".in.pfd")
        for oname, (ropts, kw) in OPTSETS.items():
            tag = f"{name}.{oname}"
            out = os.path.join(HERE, tag + ".out.pfd")
            drat = os.path.join(HERE, tag + ".drat")
            mp = os.path.join(HERE, tag + ".map")
            env = dict(os.environ, PFREF_MAP_OUT=mp)
            if learnt:
                env["PFREF_LEARNT"] = learnt
            r = subprocess.run([REF, cnf_path, inp, out] + ropts + ["-proof", f"--proofout={drat}"],
                               capture_output=True, text=True, timeout=600, env=env)
            if not os.path.exists(out):
                print("skip (refuted at parse time)", tag)
                continue
            gz(out)
            with open(out + ".proofpos") as f:
                pos = int(f.read())
            os.remove(out + ".proofpos")
            with open(drat, "rb") as f:
                proof = f.read()[pos:]
            os.remove(drat)
            with open(os.path.join(HERE, tag + ".proof"), "wb") as f:
                f.write(proof)
            gz(os.path.join(HERE, tag + ".proof"))
            entry = {"name": tag, "input": name + ".in.pfd.gz", "output": tag + ".out.pfd.gz",
                     "proof": tag + ".proof.gz", "ref_options": ropts, "options": kw}
            if learnt:
                entry["learnt"] = learnt
            if os.path.exists(mp):
                gz(mp)
                entry["map"] = tag + ".map.gz"
            manifest.append(entry)
            print(tag, r.stdout.strip().splitlines()[-1][:100])
        os.remove(cnf_path)
        if os.path.exists(inp):
            gz(inp)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    main_dimacs()
    main_els()


if __name__ == "__main__":
    main()
