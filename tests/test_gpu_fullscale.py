"""GPU parity at the benchmark workloads' own shapes (not scaled down).

* Nested-loop join, 2^16 x 2^16 tuples, keys uniform in [0, 2^24): the
  per-window NLJ shape of the plain HIP benchmark (multi-split left grid).
* Batched windows: 4096 windows per launch, per-side sizes lognormal
  (median 2^14, sigma 1) clipped to [2^8, 2^20], keys Zipf theta in
  {0, 0.8, 1.2} over [0, 2^28); the 64 largest windows and 64 random ones
  are checked window by window against the oracle.
* Learned partitioning at 2^22 tuples per side on the drifting mixture
  (grouped count + C1, with peeled heavy keys), in the training epoch and
  640 windows later (a stale model).
* Eager sliding window, W = 2^22 per side, events of Poisson(2^16) tuples
  per side, Zipf 1.0 keys over [0, 2^26): count + C1 after the event that
  fills the window and after every 64th event from there, against the
  oracle's window join; the serving loop must give the push results.

Bar: bit-exact count, C1 and (where pairs are enumerated) C2.
"""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu

NLJ, HJ, SMJ = 1, 2, 3
F_C2, F_GROUPED, F_LEARNED = 1, 2, 4


def test_nlj_window16_uniform24(ctx, torch_mod, orc, sjbmod):
    n = 1 << 16
    for w in range(2):
        print(f"case nlj window16 w={w}", flush=True)
        R = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=123 + 2 * w, ts_base=w * n)
        S = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=456 + 2 * w, ts_base=w * n)
        o, _ = orc.lazy_hj(R, S, 1 << 16)
        st = ctx.lazy_join(NLJ, 0, F_C2, to_dev(torch_mod, R), n, to_dev(torch_mod, S), n)
        assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
        assert 150 < o.count < 400  # ~2^32 / 2^24 = 256 matches


def lognormal_sizes(rng, n_win):
    sz = np.exp(rng.normal(np.log(2.0 ** 14), 1.0, size=n_win))
    return np.clip(np.round(sz), 2 ** 8, 2 ** 20).astype(np.uint64)


@pytest.mark.parametrize("theta", [0.0, 0.8, 1.2])
def test_batch_4096_lognormal_windows(ctx, torch_mod, orc, sjbmod, theta):
    print(f"case batch4096 theta={theta}", flush=True)
    rng = np.random.default_rng(int(theta * 10) + 77)
    n_win = 4096
    nl, nr = lognormal_sizes(rng, n_win), lognormal_sizes(rng, n_win)
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    dist = sjbmod.DIST_ZIPF if theta > 0 else sjbmod.DIST_UNIFORM
    R = sjbmod.generate(int(lo[-1]), dist, 1 << 28, theta, seed=1000 + int(theta * 10))
    S = sjbmod.generate(int(ro[-1]), dist, 1 << 28, theta, seed=2000 + int(theta * 10))
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    order = np.argsort(nl + nr)
    pick = np.unique(np.concatenate([order[-64:], rng.choice(n_win, 64, replace=False)]))
    modes = [F_GROUPED] if theta > 1.0 else [F_C2, F_GROUPED]
    ref = {}
    for w in pick:
        r, s = R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]]
        o, _ = orc.lazy_smj(r, s)
        ref[(w, F_GROUPED)] = (o.count, o.c1)
        if F_C2 in modes:
            D = 1 << max(10, int(len(r)).bit_length())
            o2, _ = orc.lazy_hj(r, s, D)
            assert (o2.count, o2.c1) == (o.count, o.c1)
            ref[(w, F_C2)] = o2.triple()
    for alg in (HJ, SMJ):
        for flags in modes:
            print(f"  alg={alg} flags={flags}", flush=True)
            res = ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro)
            for w in pick:
                x = res[w]
                if flags == F_GROUPED:
                    got = (int(x["count"]), int(x["checksum1"]))
                else:
                    got = (int(x["count"]), int(x["checksum1"]), int(x["checksum2"]))
                assert got == ref[(w, flags)], (alg, flags, int(w))
    # totals over every window: HJ == SMJ (grouped)
    a = ctx.lazy_join_batch(HJ, F_GROUPED, dR, lo, dS, ro)
    b = ctx.lazy_join_batch(SMJ, F_GROUPED, dR, lo, dS, ro)
    assert np.array_equal(a[["count", "checksum1"]], b[["count", "checksum1"]])


@pytest.fixture(scope="module")
def fitted_full(sjbmod, torch_mod):
    """MLP-root model fitted on windows 0..3 (2^20 tuples each) of the mixture."""
    train = np.concatenate([sjbmod.generate_mix(1 << 20, w, seed=70 + w) for w in range(4)])
    d_train = to_dev(torch_mod, train)
    c = sjbmod.Context(max_window=1 << 22, distinct_keys=1 << 16)
    rep = c.rmi_fit(d_train, len(train), root=sjbmod.RMI_MLP, steps=400, lr=0.01)
    assert rep.fallback == 0
    yield c
    c.close()


@pytest.mark.parametrize("window", [4, 640])
def test_learned_window22(fitted_full, torch_mod, orc, sjbmod, window):
    c = fitted_full
    n = 1 << 22
    print(f"case learned window22 window={window}", flush=True)
    R = sjbmod.generate_mix(n, window, seed=971 + window)
    S = sjbmod.generate_mix(n, window, seed=972 + window)
    o, _ = orc.lazy_smj(R, S)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for alg in (HJ, SMJ):
        for flags in (F_GROUPED | F_LEARNED, F_GROUPED):
            st = c.lazy_join(alg, 0, flags, dR, n, dS, n)
            assert (st.result_count, st.checksum1) == (o.count, o.c1), (alg, flags)
    rep = c.partition_report(HJ, F_GROUPED | F_LEARNED, dR, n, dS, n)
    assert rep.heavy_left > 0 and rep.heavy_right > 0


def test_slide_window22_poisson16(ctx, orc, sjbmod, torch_mod):
    W, mean, n_events = 1 << 22, 1 << 16, 256
    mb = 1 << 17
    rng = np.random.default_rng(55)
    nl = np.minimum(rng.poisson(mean, n_events), mb)
    nr = np.minimum(rng.poisson(mean, n_events), mb)
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    L = sjbmod.generate(int(lo[-1]), sjbmod.DIST_ZIPF, 1 << 26, 1.0, seed=501)
    R = sjbmod.generate(int(ro[-1]), sjbmod.DIST_ZIPF, 1 << 26, 1.0, seed=502)
    full = int(np.argmax((lo[1:] >= W) & (ro[1:] >= W)))
    assert lo[full + 1] >= W and ro[full + 1] >= W
    check = [i for i in range(full, n_events, 64)] + [n_events - 1]
    print(f"case slide window22 fill_event={full} checks={check}", flush=True)
    ref = orc.sliding_var(L, lo[1:][check], R, ro[1:][check], W)
    sl = ctx.slide(W, mb)
    got = [sl.push(L[lo[i]:lo[i + 1]], R[ro[i]:ro[i + 1]]) for i in range(n_events)]
    assert [got[i] for i in check] == ref
    assert ref[0][0] > 0
    info = sl.info()
    assert info.events == n_events and info.tuples_left == lo[-1]
    # the serving loop over the same events (pinned batches, all arrived)
    Lp = torch_mod.from_numpy(L.view(np.uint8)).pin_memory()
    Rp = torch_mod.from_numpy(R.view(np.uint8)).pin_memory()
    sl.reset()
    res, lat, _ = sl.serve(Lp, lo, Rp, ro, np.zeros(n_events))
    assert [(int(x["count"]), int(x["checksum1"])) for x in res] == got
    assert np.all(lat > 0)
    sl.destroy()
