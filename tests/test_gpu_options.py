"""The option surface of the reference's GPU build (src/gpu/options.cpp:24-43,
src/gpu/options.cu:36-60) through the C ABI, and parity of the options that
change results (electionsmax, lcveclausemax, aggresivesort) with the oracle;
the occurrence table built by one atomic per occurrence (otpranges=0: the
build past 12,288 literal ranges) and with other tiles, and the election with
2, 4 or 8 lanes per candidate and other LDS reservations, also match it."""
import numpy as np
import pytest

from oracle import oracle
from parafrost_amd import gen
from tests.common import diff, oracle_kw

pytestmark = pytest.mark.gpu

# every option name of the GPU build's simplifier, with its default value
GPU_OPTIONS = {
    "aggresivesort": 0, "all": 0, "bce": 0, "ere": 1, "sub": 1, "solve": 1, "ve": 1, "veextend": 1,
    "lcvefast": 0, "electionsmin": 2, "electionsmax": 3000, "lcveclausemax": 30000, "eliminatedlitsmin": 500,
    "mupos": 32, "muneg": 32, "phases": 5, "collectfreq": 2, "nstreams": 6, "literalsmul": 1.0,
    "syncalways": 0, "profilegpu": 0, "veminthreads": 4, "subminthreads": 4, "ereminthreads": 4,
    "veminblocks": 0.5, "subminblocks": 0.5, "ereminblocks": 0.5, "vefunction": 0, "velitsbound": 0,
    "bcemaxoccurs": 3000, "submaxoccurs": 3000, "subclausemax": 100, "ereextend": 1, "eremaxoccurs": 3000,
    "ereclausemax": 250, "resolventmax": 100, "xormaxarity": 10,
}


def test_gpu_option_names_accepted():
    from parafrost_amd.simplifier import Simplifier
    s = Simplifier(0)
    try:
        for k, v in GPU_OPTIONS.items():
            s.set_option(k, v)
    finally:
        s.close()


@pytest.mark.parametrize("name", ["lcvefast", "vefunction"])
def test_unprovided_heuristics_refused(name):
    from parafrost_amd.simplifier import PfsError, Simplifier
    s = Simplifier(0)
    try:
        s.set_option(name, 0)
        with pytest.raises(PfsError, match="BAD_INPUT.*not provided"):
            s.set_option(name, 1)
    finally:
        s.close()


@pytest.mark.parametrize("name,value", [("nstreams", 5), ("nstreams", 33), ("literalsmul", 5),
                                        ("veminblocks", 2), ("electionsmax", 0), ("ereextend", 4)])
def test_out_of_range_refused(name, value):
    from parafrost_amd.simplifier import PfsError, Simplifier
    s = Simplifier(0)
    try:
        with pytest.raises(PfsError, match="BAD_INPUT"):
            s.set_option(name, value)
    finally:
        s.close()


FORMULAS = {
    "uniform3": lambda: gen.uniform_ksat(3000, 12780, 3, seed=901),
    "heavytail": lambda: gen.heavytail(20000, 80000, seed=902),
    "tseitin": lambda: gen.prepare(gen.tseitin(20000, 400, 40, seed=903))[0],
}
LIMITS = [{"electionsmax": 8}, {"electionsmax": 40}, {"lcveclausemax": 2}, {"lcveclausemax": 3},
          {"aggresivesort": 1}, {"aggresivesort": 1, "all": 1, "phases": 3}, {"literalsmul": 0},
          {"syncalways": 1}, {"otpranges": 0}, {"otptile": 10, "otpxcd": 0}, {"mislanes": 2},
          {"mislanes": 8, "mislds": 0}, {"mislanes": 4, "mislds": 20480, "autotune": 0}]


@pytest.mark.parametrize("family", sorted(FORMULAS))
@pytest.mark.parametrize("opt", range(len(LIMITS)))
def test_options_match_oracle(family, opt):
    from parafrost_amd.simplifier import Simplifier
    cnf = FORMULAS[family]()
    o = dict(LIMITS[opt])
    print("options", family, o, flush=True)
    s = Simplifier(0, phases=1, bce=1)
    try:
        for k, v in o.items():
            s.set_option(k, v)
        s.upload(cnf.maxvar, cnf.offsets, cnf.lits, trail=cnf.trail)
        g = s.result(s.simplify())
    finally:
        s.close()
    kw = {"phases": 1, "bce_en": 1}
    # launch shapes and stream behaviour have no oracle counterpart (results do not depend on them)
    shaping = ("literalsmul", "syncalways", "otpranges", "otptile", "otpxcd", "mislanes", "mislds", "autotune")
    kw.update(oracle_kw({k: v for k, v in o.items() if k not in shaping}))
    vs = np.zeros(cnf.maxvar + 1, np.uint8)
    vs[np.asarray(cnf.trail, np.int64) >> 1] = 2  # root-assigned variables are frozen
    r = oracle.simplify(cnf.maxvar, cnf.offsets, cnf.lits, nassigned=int(cnf.trail.shape[0]), det=1, vstate=vs,
                        **kw)
    d = diff(g, r)
    assert not d, d
    if g.rounds and r.elected.shape[0]:
        assert g.rounds[0]["elected"] == int(r.elected[0])
