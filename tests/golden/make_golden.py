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
with.  Run from the repository root:  python tests/golden/make_golden.py
"""
import gzip
import json
import os
import subprocess
import sys

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
    yield "u3_n60", gen.uniform_ksat(60, 256, 3, seed=11)
    yield "u3_n300", gen.uniform_ksat(300, 1278, 3, seed=12)
    yield "u3_n2000", gen.uniform_ksat(2000, 8520, 3, seed=13)
    yield "u3lo_n1000", gen.uniform_ksat(1000, 3000, 3, seed=14)
    yield "u2_n400", gen.uniform_ksat(400, 500, 2, seed=15)
    yield "ts_g300", gen.tseitin(300, 20, 3, seed=21)
    yield "ts_g2000", gen.tseitin(2000, 60, 10, seed=22)
    yield "ts_g4000", gen.tseitin(4000, 100, 12, seed=23)
    yield "ht_v300", gen.heavytail(300, 1200, seed=31)
    yield "ht_v2000", gen.heavytail(2000, 8000, seed=32)
    yield "ht_v3000", gen.heavytail(3000, 12000, seed=33)


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
    for name, cnf in cases():
        cnf_path = os.path.join(HERE, name + ".cnf")
        gen.write_dimacs(cnf, cnf_path)
        inp = os.path.join(HERE, name + ".in.pfd")
        for oname, (ropts, kw) in OPTSETS.items():
            tag = f"{name}.{oname}"
            out = os.path.join(HERE, tag + ".out.pfd")
            drat = os.path.join(HERE, tag + ".drat")
            mp = os.path.join(HERE, tag + ".map")
            env = dict(os.environ, PFREF_MAP_OUT=mp)
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


if __name__ == "__main__":
    main()
