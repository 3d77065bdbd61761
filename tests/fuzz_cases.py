"""Seeded small dense random formulas that exercise units found by
strengthening, cross-variable ERE deletions, refutations, gate detection and
learnt clauses (shared by the CPU and GPU fuzz tests)."""
import random

import numpy as np

from parafrost_amd import gen

# (reference-name options, oracle keyword options)
OPTION_SETS = [
    ({}, {}),
    ({"phases": 1, "bce": 1}, {"phases": 1, "bce_en": 1}),
    ({"all": 1}, {"all_en": 1}),
    ({"all": 1, "mupos": 10, "muneg": 10, "phases": 3}, {"all_en": 1, "mu_pos": 10, "mu_neg": 10, "phases": 3}),
    ({"sub": 0, "ve+": 0}, {"sub_en": 0, "ve_plus_en": 0}),
]


def case(i):
    """Returns (cnf, flags, lbd, ref_options, oracle_options) for case i."""
    rng = random.Random(0xF022 + i)
    V = rng.randint(5, 80)
    C = rng.randint(V, 6 * V)
    lens = rng.choice([[2], [2, 3], [2, 2, 3, 4], [3], [2, 3, 3, 5, 7]])
    dup = rng.choice([0.0, 0.05, 0.3])
    learnt = rng.choice([0.0, 0.0, 0.3])
    cls = []
    for _ in range(C):
        if cls and rng.random() < dup:
            c = list(rng.choice(cls))
            rng.shuffle(c)
            cls.append(c)
            continue
        k = min(rng.choice(lens), V)
        vs = rng.sample(range(1, V + 1), k)
        cls.append([v if rng.random() < 0.5 else -v for v in vs])
    cnf = gen.from_clauses(V, cls)
    flags = lbd = None
    if learnt:
        flags = np.array([1 if rng.random() < learnt else 0 for _ in cls], np.uint8)
        lbd = np.array([rng.randint(1, 8) for _ in cls], np.uint32)
    ro, oo = OPTION_SETS[i % len(OPTION_SETS)]
    return cnf, flags, lbd, ro, oo
