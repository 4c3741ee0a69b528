"""Loading the committed golden fixtures of tests/golden/ (written and
cross-checked by tests/golden/make_fixtures.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WINDOWS = ["window12_uniform", "window12_zipf"]


def load_window(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    triple = tuple(int(x) for x in z["triple"])
    return z["R"], z["S"], triple, z["pairs"]


def load_cfg1():
    with open(os.path.join(GOLDEN, "cfg1.json")) as f:
        return json.load(f)


def rows_sorted(keys, rvals, svals):
    """(key, left val, right val) rows in lexicographic order."""
    rows = np.stack([keys, rvals, svals], axis=1).astype(np.uint32)
    return rows[np.lexsort((rows[:, 2], rows[:, 1], rows[:, 0]))]
