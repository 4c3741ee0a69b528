"""The oracle and the input generator against the committed golden fixtures
(tests/golden/): two small windows with their full sorted pair lists, and
the reference's CPU-runnable configuration (cfg1: 16 windows of 2^16 per
side, keys uniform in [0, 2^16)) as input digests plus every window's
count / C1 / C2.  The fixtures were cross-checked three ways when written
(oracle HJ / SMJ / NLJ, numpy pair enumeration, per-key closed forms); here
they pin the oracle and the generator against later changes."""
import hashlib

import numpy as np
import pytest

from golden_io import WINDOWS, load_cfg1, load_window, rows_sorted
from helpers import np_join


@pytest.mark.parametrize("name", WINDOWS)
def test_window_fixture(orc, name):
    R, S, triple, pairs = load_window(name)
    assert np_join(R, S) == triple
    o_hj, _ = orc.lazy_hj(R, S, 1 << 12)
    o_smj, _ = orc.lazy_smj(R, S, want_c2=True)
    o_nlj, p = orc.lazy_nlj(R, S, want_pairs=True, cap=max(triple[0], 1))
    assert o_hj.triple() == o_smj.triple() == o_nlj.triple() == triple
    assert np.array_equal(rows_sorted(p["lkey"], p["lval"], p["rval"]), pairs)


def test_cfg1_fixture(orc, sjbmod):
    doc = load_cfg1()
    g = doc["generator"]
    R = sjbmod.generate(g["n"], sjbmod.DIST_UNIFORM, g["key_range"], 0.0, seed=g["seed_left"])
    S = sjbmod.generate(g["n"], sjbmod.DIST_UNIFORM, g["key_range"], 0.0, seed=g["seed_right"])
    assert hashlib.sha256(R.tobytes()).hexdigest() == doc["sha256_left"]
    assert hashlib.sha256(S.tobytes()).hexdigest() == doc["sha256_right"]
    W = doc["window"]
    for w, ref in enumerate(doc["windows"]):
        o, _ = orc.lazy_hj(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W], g["key_range"])
        assert o.triple() == (ref["count"], ref["c1"], ref["c2"]), w
