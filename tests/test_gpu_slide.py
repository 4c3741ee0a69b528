"""GPU parity of the eager sliding-window join (sjb_slide_*): after every
event (one left and one right batch of Poisson-distributed size) the window
is the last W tuples of each side; count and C1 of its join must equal the
oracle's brute-force window join (or_sliding_var) bit for bit, through the
synchronous push and through the serving loop, across table rebuilds, with
key 0xFFFFFFFF and empty batches; state admission obeys the HBM budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def events(sjbmod, n_events, mean, max_batch, dist, key_range, theta, seed):
    rng = np.random.default_rng(seed)
    nl = np.minimum(rng.poisson(mean, n_events), max_batch)
    nr = np.minimum(rng.poisson(mean, n_events), max_batch)
    nl[3] = 0
    nr[5] = 0
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    L = sjbmod.generate(int(lo[-1]), dist, key_range, theta, seed=seed + 1)
    R = sjbmod.generate(int(ro[-1]), dist, key_range, theta, seed=seed + 2)
    return L, lo, R, ro


def test_slide_push_matches_oracle(ctx, orc, sjbmod):
    W, mb = 4096, 1024
    L, lo, R, ro = events(sjbmod, 60, 512, mb, sjbmod.DIST_ZIPF, 1 << 16, 1.0, 5)
    L["key"][::97] = 0xFFFFFFFF
    R["key"][::89] = 0xFFFFFFFF
    ref = orc.sliding_var(L, lo[1:], R, ro[1:], W)
    sl = ctx.slide(W, mb)
    got = [sl.push(L[lo[i]:lo[i + 1]], R[ro[i]:ro[i + 1]]) for i in range(len(lo) - 1)]
    assert got == ref
    assert ref[-1][0] > 0
    info = sl.info()
    assert info.events == 60 and info.tuples_left == lo[-1]
    sl.destroy()


def test_serve_matches_push(ctx, orc, sjbmod, torch_mod):
    W, mb = 8192, 2048
    L, lo, R, ro = events(sjbmod, 80, 1024, mb, sjbmod.DIST_ZIPF, 1 << 18, 1.0, 9)
    ref = orc.sliding_var(L, lo[1:], R, ro[1:], W)
    # pinned host copies for asynchronous transfers
    Lp = torch_mod.from_numpy(L.view(np.uint8)).pin_memory()
    Rp = torch_mod.from_numpy(R.view(np.uint8)).pin_memory()
    sl = ctx.slide(W, mb)
    arrival = np.arange(80) * 20e-6
    res, lat, el = sl.serve(Lp, lo, Rp, ro, arrival)
    assert [(int(x["count"]), int(x["checksum1"])) for x in res] == ref
    assert np.all(lat > 0) and el > arrival[-1]
    sl.reset()
    res2, _, _ = sl.serve(Lp, lo, Rp, ro, np.zeros(80))
    assert np.array_equal(res, res2)
    sl.destroy()


def test_slide_rebuilds(ctx, orc, sjbmod):
    # all-distinct keys: removed keys leave empty-count slots, forcing rebuilds
    W, mb = 2048, 512
    L, lo, R, ro = events(sjbmod, 120, 400, mb, sjbmod.DIST_UNIFORM, 1 << 12, 0.0, 13)
    L["key"][1::2] = np.arange(1, len(L), 2, dtype=np.uint32) * 7919 + (1 << 20)
    ref = orc.sliding_var(L, lo[1:], R, ro[1:], W)
    sl = ctx.slide(W, mb)
    got = [sl.push(L[lo[i]:lo[i + 1]], R[ro[i]:ro[i + 1]]) for i in range(len(lo) - 1)]
    assert got == ref
    assert sl.info().rebuilds >= 3
    sl.destroy()


def test_budget_admission(sjbmod):
    W, mb = 1 << 16, 1 << 12
    with sjbmod.Context(max_window=1 << 16) as probe:
        per = probe.slide(W, mb).info().device_bytes
        base = probe.device_bytes()
    with sjbmod.Context(max_window=1 << 16, hbm_budget=base + int(2.5 * per)) as c:
        a = c.slide(W, mb)
        b = c.slide(W, mb)
        with pytest.raises(sjbmod.SjbError) as e:
            c.slide(W, mb)
        assert e.value.code == -7
        b.destroy()
        d = c.slide(W, mb)  # freed state is admitted again
        assert d.info().device_bytes == per
        a.destroy()
        d.destroy()
