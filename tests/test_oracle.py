"""Pins the CPU oracle (no GPU needed).

Each oracle function is checked against something it does not share code
with: published SplitMix64 outputs, brute force (NLJ) on tiny inputs, the
per-key closed forms (SMJ), an independent numpy restatement, and special
cases whose answers the mathematics fixes.
"""
import numpy as np
import pytest

from helpers import make, np_join, np_pair_rows, sort_rows, zipf_keys, np_mix64


def test_mix64_published_vectors(orc):
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): the
    # generator state advances by 0x9E3779B97F4A7C15 and outputs mix64(state).
    # Seeded with 0 its first output is 0xE220A8397B1DCDAF.
    L = orc.lib()
    assert L.or_mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    assert L.or_mix64(0) == 0
    state = 1234567
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423,
              4593380528125082431, 16408922859458223821]
    got = []
    for _ in expect:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        got.append(L.or_mix64(state))
    assert got == expect


def test_numpy_restatement_matches_c(orc):
    rng = np.random.default_rng(7)
    xs = rng.integers(0, 2**63, 1000, dtype=np.uint64)
    L = orc.lib()
    assert [int(v) for v in np_mix64(xs)] == [L.or_mix64(int(x)) for x in xs]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("key_range", [4, 37, 1000, 1 << 20])
def test_nlj_hj_smj_agree_with_numpy(orc, seed, key_range):
    rng = np.random.default_rng(seed)
    nL, nR = int(rng.integers(0, 300)), int(rng.integers(0, 300))
    R, S = make(nL, key_range, seed), make(nR, key_range, seed + 100)
    ref = np_join(R, S)
    nlj, pn = orc.lazy_nlj(R, S, want_pairs=True)
    for D in (1, 7, 1000, 1 << 16):
        hj, ph = orc.lazy_hj(R, S, D, want_c2=True, want_pairs=True)
        assert hj.triple() == ref
        assert np.array_equal(sort_rows(ph), sort_rows(pn))
    smj, ps = orc.lazy_smj(R, S, want_c2=True, want_pairs=True)
    smj_closed, _ = orc.lazy_smj(R, S)
    assert nlj.triple() == ref
    assert smj.triple() == ref
    # closed form: count and C1 from per-key sums, no pair enumeration
    assert (smj_closed.count, smj_closed.c1) == ref[:2]
    assert np.array_equal(sort_rows(pn), sort_rows(np_pair_rows(R, S)))
    assert np.array_equal(sort_rows(ps), sort_rows(pn))


def test_special_cases(orc):
    R = make(50, 1, 1)  # every key 0
    S = make(70, 1, 2)
    r, _ = orc.lazy_hj(R, S, 1000)
    assert r.count == 50 * 70
    # all-equal keys: C1 = (sum A)(sum B) — one key group
    from helpers import np_hash_a, np_hash_b
    with np.errstate(over="ignore"):
        c1 = int(np.sum(np_hash_a(R["key"], R["val"]), dtype=np.uint64) *
                 np.sum(np_hash_b(S["key"], S["val"]), dtype=np.uint64))
    assert r.c1 == c1
    # disjoint key sets
    R2 = make(100, 10, 3)
    S2 = make(100, 10, 4)
    S2["key"] += 10
    for fn in (lambda: orc.lazy_hj(R2, S2, 7)[0], lambda: orc.lazy_smj(R2, S2, True)[0],
               lambda: orc.lazy_nlj(R2, S2)[0]):
        assert fn().triple() == (0, 0, 0)
    # empty sides
    E = make(0, 10, 5)
    assert orc.lazy_hj(E, S2, 7)[0].triple() == (0, 0, 0)
    assert orc.lazy_smj(R2, E, True)[0].triple() == (0, 0, 0)
    # max key value is an ordinary key
    R3 = make(10, 1, 6)
    S3 = make(10, 1, 7)
    R3["key"] = 0xFFFFFFFF
    S3["key"] = 0xFFFFFFFF
    assert orc.lazy_smj(R3, S3, True)[0].count == 100
    assert orc.lazy_hj(R3, S3, 1 << 16)[0].count == 100


def test_skewed_closed_form(orc):
    keys_r = zipf_keys(5000, 1.2, 1 << 12, 1)
    keys_s = zipf_keys(5000, 1.2, 1 << 12, 2)
    R = make(5000, 1, 11, keys=keys_r)
    S = make(5000, 1, 12, keys=keys_s)
    ref = np_join(R, S)
    smj, _ = orc.lazy_smj(R, S, want_c2=True)
    assert smj.triple() == ref
    closed, _ = orc.lazy_smj(R, S)
    assert (closed.count, closed.c1) == ref[:2]


def test_expected_count_uniform(orc):
    # Uniform keys over K: E[count] = nL*nR/K, Var ~= nL*nR/K.
    n, K = 1 << 15, 1 << 14
    R, S = make(n, K, 21), make(n, K, 22)
    r, _ = orc.lazy_hj(R, S, K)
    mu = n * n / K
    assert abs(r.count - mu) < 6 * np.sqrt(mu * 3)


@pytest.mark.parametrize("B", [16, 33])
def test_eager_equals_lazy(orc, B):
    # canonical windows: all batches of a window pushed L0,R0,L1,R1,...;
    # per-batch counts (build own side, probe the other) sum to the lazy count.
    n = 8 * B
    R, S = make(n, 50, 31), make(n, 50, 32)
    e = orc.Eager(1000, n)
    tot = [0, 0, 0]
    for b in range(8):
        for side, src in ((1, R), (0, S)):
            res = e.execute(src[b * B:(b + 1) * B], side)
            tot = [(t + v) % 2**64 for t, v in zip(tot, res.triple())]
    assert tuple(tot) == np_join(R, S)
    e.clear()
    assert e.execute(R[:B], 1).count == 0


def test_sliding_bruteforce(orc):
    n, W, slide = 600, 200, 50
    R, S = make(n, 40, 41), make(n, 40, 42)
    res = orc.sliding(R, S, W, slide)
    for t, (c, c1, _) in enumerate(res):
        end = (t + 1) * slide
        beg = max(0, end - W)
        ref = np_join(R[beg:end], S[beg:end], with_c2=False)
        assert (c, c1) == ref[:2]


def test_cpu_baseline_runs(orc):
    n, W = 1 << 14, 1 << 12
    R, S = make(n, 1 << 12, 51), make(n, 1 << 12, 52)
    secs, tot = orc.bench_lazy_hj(R, S, W, 1 << 12, 2)
    assert secs >= 0
    exp = [0, 0, 0]
    for w in range(0, n, W):
        r, _ = orc.lazy_hj(R[w:w + W], S[w:w + W], 1 << 12)
        exp = [(a + b) % 2**64 for a, b in zip(exp, r.triple())]
    assert tot.triple() == tuple(exp)


# ---- learned partitioner root: the float64 MLP gradient ------------------
def _mlp_case(seed, n):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 0.5, 1153)
    x = rng.uniform(0, 1, n)
    t = np.sort(rng.uniform(0, 1, n))
    return P, x, t


def test_mlp_oracle_gradient_finite_differences(orc):
    # central differences pin the analytic back-propagation of the oracle
    P, x, t = _mlp_case(1, 50)
    loss, g = orc.mlp_grad(P, x, t)
    assert abs(loss - orc.mlp_loss(P, x, t)) < 1e-15
    rng = np.random.default_rng(2)
    idx = np.concatenate([np.arange(0, 64), rng.choice(np.arange(64, 1153), 120, replace=False)])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (orc.mlp_loss(Pp, x, t) - orc.mlp_loss(Pm, x, t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_oracle_special_cases(orc):
    # all weights zero except b3: y = b3, d loss / d b3 = 2 mean(b3 - t)
    P = np.zeros(1153)
    P[1152] = 0.25
    x = np.linspace(0, 1, 11)
    t = np.linspace(0, 1, 11)
    loss, g = orc.mlp_grad(P, x, t)
    assert abs(loss - np.mean((0.25 - t) ** 2)) < 1e-15
    assert abs(g[1152] - 2 * np.mean(0.25 - t)) < 1e-15
    assert np.all(g[:1152] == 0)  # every hidden unit is dead (relu(0) = 0, relu'(0) = 0)
