"""Input builders and an independent numpy restatement of the checksums.

The numpy code below recomputes count/C1/C2 from an explicit pair list built
with numpy (argsort + searchsorted), using uint64 wrap-around arithmetic; it
shares no code with the C oracle, so agreement pins the oracle's arithmetic.
"""
import numpy as np

TUPLE_DTYPE = np.dtype([("key", "<u4"), ("val", "<u4"), ("ts", "<u8")])
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SEED_A = np.uint64(0x243F6A8885A308D3)
SEED_B = np.uint64(0x13198A2E03707344)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def np_mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def np_hash_a(k, v):
    return np_mix64(((np.asarray(k, np.uint64) << np.uint64(32)) | np.asarray(v, np.uint64)) ^ SEED_A)


def np_hash_b(k, v):
    return np_mix64(((np.asarray(k, np.uint64) << np.uint64(32)) | np.asarray(v, np.uint64)) ^ SEED_B)


def np_pair_hash(k, rv, sv):
    a = np_hash_a(k, rv)
    with np.errstate(over="ignore"):
        return np_mix64(a + GOLDEN * ((np.asarray(sv, np.uint64) << np.uint64(1)) | np.uint64(1)))


def np_pairs(R, S):
    """All (i, j) with R[i].key == S[j].key, via sorting (numpy only)."""
    order = np.argsort(R["key"], kind="stable")
    rk = R["key"][order]
    lo = np.searchsorted(rk, S["key"], side="left")
    hi = np.searchsorted(rk, S["key"], side="right")
    cnt = hi - lo
    js = np.repeat(np.arange(len(S)), cnt)
    starts = np.repeat(lo, cnt)
    offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    iis = order[starts + offs]
    return iis, js


def np_join(R, S, with_c2=True):
    i, j = np_pairs(R, S)
    k = R["key"][i]
    with np.errstate(over="ignore"):
        c1 = int(np.sum(np_hash_a(k, R["val"][i]) * np_hash_b(k, S["val"][j]), dtype=np.uint64))
        c2 = int(np.sum(np_pair_hash(k, R["val"][i], S["val"][j]), dtype=np.uint64)) if with_c2 else 0
    return len(i), c1, c2


def np_pair_rows(R, S):
    i, j = np_pairs(R, S)
    out = np.zeros(len(i), dtype=[("lkey", "<u4"), ("rkey", "<u4"), ("lval", "<u4"),
                                  ("rval", "<u4"), ("lts", "<u8"), ("rts", "<u8")])
    out["lkey"] = R["key"][i]
    out["rkey"] = S["key"][j]
    out["lval"] = R["val"][i]
    out["rval"] = S["val"][j]
    out["lts"] = R["ts"][i]
    out["rts"] = S["ts"][j]
    return out


def make(n, key_range, seed, ts_base=0, keys=None):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=TUPLE_DTYPE)
    a["key"] = rng.integers(0, key_range, n, dtype=np.uint64).astype(np.uint32) if keys is None else keys
    a["val"] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    a["ts"] = np.arange(ts_base, ts_base + n, dtype=np.uint64)
    return a


def zipf_keys(n, theta, key_range, seed):
    """Zipf-like skewed keys for tests (numpy; not the product generator)."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, key_range + 1, dtype=np.float64)
    p = ranks ** (-theta)
    p /= p.sum()
    r = rng.choice(key_range, size=n, p=p)
    perm = rng.permutation(key_range)
    return perm[r].astype(np.uint32)


def sort_rows(p):
    return np.sort(p, order=["lkey", "lval", "rval", "lts", "rts", "rkey"])
