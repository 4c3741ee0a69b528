"""Writes the golden fixtures of tests/golden/ (run from the repository root:
`python tests/golden/make_fixtures.py`; needs the built oracle and
libsjb_hip.so's host-side generator, no GPU).

* window12_uniform.npz / window12_zipf.npz: one small window each (inputs
  stored), its count / C1 / C2 and the sorted (r.key, r.val, s.val) pair
  list.
* cfg1.json: the reference's CPU-runnable configuration (two streams of
  2^20 tuples, keys uniform in [0, 2^16), 16 tumbling windows of 2^16 per
  side; inputs from the build's generator, stored as SHA-256 digests) with
  every window's count / C1 / C2.

Every expected value is computed three ways before it is written — the C
oracle (HJ, SMJ and NLJ restatements), the independent numpy pair
enumeration of tests/helpers.py, and the per-key closed forms count =
sum_k L_k R_k and C1 = sum_k (sum A)(sum B) — and the script stops if any
two disagree.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "gpu-stream-join-benchmark_amd", "py"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
import sjb  # noqa: E402
from helpers import np_hash_a, np_hash_b, np_join, np_pairs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def closed_form(R, S):
    """count and C1 per key group (no pair enumeration)."""
    keys = np.union1d(R["key"], S["key"])
    with np.errstate(over="ignore"):
        a = np_hash_a(R["key"], R["val"])
        b = np_hash_b(S["key"], S["val"])
    ri = np.searchsorted(keys, R["key"])
    si = np.searchsorted(keys, S["key"])
    cnt_l = np.bincount(ri, minlength=len(keys)).astype(np.uint64)
    cnt_r = np.bincount(si, minlength=len(keys)).astype(np.uint64)
    sum_a = np.zeros(len(keys), np.uint64)
    sum_b = np.zeros(len(keys), np.uint64)
    np.add.at(sum_a, ri, a)
    np.add.at(sum_b, si, b)
    with np.errstate(over="ignore"):
        count = int(np.sum(cnt_l * cnt_r, dtype=np.uint64))
        c1 = int(np.sum(sum_a * sum_b, dtype=np.uint64))
    return count, c1


def sorted_pairs(R, S):
    i, j = np_pairs(R, S)
    return rows_sorted(R["key"][i], R["val"][i], S["val"][j])


def rows_sorted(keys, rvals, svals):
    rows = np.stack([keys, rvals, svals], axis=1).astype(np.uint32)
    return rows[np.lexsort((rows[:, 2], rows[:, 1], rows[:, 0]))]


def window_fixture(name, R, S):
    ref = np_join(R, S)
    o_hj, _ = oracle.lazy_hj(R, S, 1 << 12)
    o_smj, _ = oracle.lazy_smj(R, S, want_c2=True)
    o_nlj, p_nlj = oracle.lazy_nlj(R, S, want_pairs=True, cap=max(ref[0], 1))
    cf = closed_form(R, S)
    assert o_hj.triple() == o_smj.triple() == o_nlj.triple() == ref, name
    assert cf == ref[:2], name
    pairs = sorted_pairs(R, S)
    assert len(pairs) == ref[0]
    got = rows_sorted(p_nlj["lkey"], p_nlj["lval"], p_nlj["rval"])
    assert np.array_equal(got, pairs), name
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), R=R, S=S, pairs=pairs,
                        triple=np.array(ref, dtype=np.uint64))
    print(f"{name}: count {ref[0]}, C1 {ref[1]:#x}, C2 {ref[2]:#x}, {len(pairs)} pairs")


def main():
    oracle.lib()
    n = 1 << 12
    R = sjb.generate(n, sjb.DIST_UNIFORM, 1 << 12, 0.0, seed=7)
    S = sjb.generate(n, sjb.DIST_UNIFORM, 1 << 12, 0.0, seed=8)
    window_fixture("window12_uniform", R, S)
    m = 1 << 10
    R = sjb.generate(m, sjb.DIST_ZIPF, 1 << 12, 1.0, seed=9)
    S = sjb.generate(m, sjb.DIST_ZIPF, 1 << 12, 1.0, seed=10)
    window_fixture("window12_zipf", R, S)

    # cfg1: 16 windows of 2^16 per side over 2^20-tuple streams
    N, W, K = 1 << 20, 1 << 16, 1 << 16
    R = sjb.generate(N, sjb.DIST_UNIFORM, K, 0.0, seed=123)
    S = sjb.generate(N, sjb.DIST_UNIFORM, K, 0.0, seed=456)
    windows = []
    for w in range(N // W):
        r, s = R[w * W:(w + 1) * W], S[w * W:(w + 1) * W]
        o, _ = oracle.lazy_hj(r, s, K)
        ref = np_join(r, s)
        assert o.triple() == ref and closed_form(r, s) == ref[:2], w
        windows.append({"count": ref[0], "c1": ref[1], "c2": ref[2]})
    doc = {
        "workload": "two streams of 2^20 tuples, keys uniform in [0, 2^16), tumbling windows of 2^16 per side",
        "generator": {"n": N, "dist": "uniform", "key_range": K, "seed_left": 123, "seed_right": 456},
        "window": W,
        "sha256_left": hashlib.sha256(R.tobytes()).hexdigest(),
        "sha256_right": hashlib.sha256(S.tobytes()).hexdigest(),
        "windows": windows,
    }
    with open(os.path.join(HERE, "cfg1.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(f"cfg1: {len(windows)} windows, total count {sum(x['count'] for x in windows)}")


if __name__ == "__main__":
    main()
