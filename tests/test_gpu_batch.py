"""GPU parity of the batched path: many uneven windows in one persistent
launch (sjb_lazy_join_batch), captured plans (hipGraph replay) and the
autotuner, against the CPU oracle window by window.

Workload shape of the batched benchmark, scaled down: window sizes
lognormal (median 2^10, sigma 1) clipped to [2^4, 2^14] per side, keys Zipf
theta in {0, 0.8, 1.2} over [0, 2^28).  Bar: per-window count, C1 (and C2
where pairs are enumerated) bit-exact.
"""
import numpy as np
import pytest

from conftest import to_dev
from helpers import make

pytestmark = pytest.mark.gpu

HJ, SMJ = 2, 3
F_C2, F_GROUPED = 1, 2


def lognormal_windows(n_win, median_log2, lo_log2, hi_log2, seed):
    rng = np.random.default_rng(seed)
    sz = np.exp(rng.normal(np.log(2.0 ** median_log2), 1.0, size=n_win))
    return np.clip(np.round(sz), 2 ** lo_log2, 2 ** hi_log2).astype(np.uint64)


def batch_case(sjbmod, n_win, theta, seed, median_log2=10, hi_log2=14):
    nl = lognormal_windows(n_win, median_log2, 4, hi_log2, seed)
    nr = lognormal_windows(n_win, median_log2, 4, hi_log2, seed + 1)
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    dist = sjbmod.DIST_ZIPF if theta > 0 else sjbmod.DIST_UNIFORM
    R = sjbmod.generate(int(lo[-1]), dist, 1 << 28, theta, seed=100 + seed)
    S = sjbmod.generate(int(ro[-1]), dist, 1 << 28, theta, seed=200 + seed)
    return R, S, lo, ro


def oracle_windows(orc, R, S, lo, ro, grouped):
    out = []
    for w in range(len(lo) - 1):
        r, s = R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]]
        if grouped:
            o, _ = orc.lazy_smj(r, s)  # closed form count / C1
            out.append((o.count, o.c1, 0))
        else:
            o, _ = orc.lazy_hj(r, s, 1 << 12)
            out.append(o.triple())
    return out


def as_triples(res, grouped):
    return [(int(x["count"]), int(x["checksum1"]), 0 if grouped else int(x["checksum2"])) for x in res]


@pytest.mark.parametrize("theta", [0.0, 0.8, 1.2])
@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_lognormal_zipf_batch(ctx, torch_mod, orc, sjbmod, alg, theta):
    print(f"case batch alg={alg} theta={theta}", flush=True)
    R, S, lo, ro = batch_case(sjbmod, 256, theta, seed=int(theta * 10) + 1)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    modes = [F_GROUPED] if theta > 1.0 else [F_C2, F_GROUPED]
    for flags in modes:
        grouped = flags == F_GROUPED
        ref = oracle_windows(orc, R, S, lo, ro, grouped)
        res = ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro)
        assert as_triples(res, grouped) == ref


def test_batch_edge_windows(ctx, torch_mod, orc):
    # empty sides, single tuples, one window of 2^18 and duplicate-heavy keys
    sizes_l = [0, 1, 5, 1 << 18, 300, 0, 77, 4096]
    sizes_r = [9, 1, 0, 1 << 17, 300, 0, 1, 4095]
    lo = np.concatenate([[0], np.cumsum(sizes_l)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(sizes_r)]).astype(np.uint64)
    R = make(int(lo[-1]), 97, 301)
    S = make(int(ro[-1]), 97, 302)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for alg in (HJ, SMJ):
        for flags in (F_C2, F_GROUPED):
            grouped = flags == F_GROUPED
            print(f"case edge alg={alg} flags={flags}", flush=True)
            res = ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro)
            assert as_triples(res, grouped) == oracle_windows(orc, R, S, lo, ro, grouped)


def test_heavy_key_large_window_grouped(ctx, torch_mod, orc, sjbmod):
    # Zipf 1.2 over 2^28 at 2^20 per side: the top key holds ~18% of each
    # side (~3.4e10 pairs); count + C1 via the aggregated build.
    n = 1 << 20
    R = sjbmod.generate(n, sjbmod.DIST_ZIPF, 1 << 28, 1.2, seed=31)
    S = sjbmod.generate(n, sjbmod.DIST_ZIPF, 1 << 28, 1.2, seed=32)
    o, _ = orc.lazy_smj(R, S)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    lo = np.array([0, n], dtype=np.uint64)
    for alg in (HJ, SMJ):
        print(f"case heavy alg={alg}", flush=True)
        st = ctx.lazy_join(alg, 0, F_GROUPED, dR, n, dS, n)
        assert (st.result_count, st.checksum1) == (o.count, o.c1)
        res = ctx.lazy_join_batch(alg, F_GROUPED, dR, lo, dS, lo)
        assert (int(res[0]["count"]), int(res[0]["checksum1"])) == (o.count, o.c1)
    assert o.count > 10 ** 10


def test_plan_graph_replay(ctx, torch_mod, sjbmod):
    R, S, lo, ro = batch_case(sjbmod, 300, 0.8, seed=7)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for alg in (HJ, SMJ):
        direct = ctx.lazy_join_batch(alg, F_C2, dR, lo, dS, ro)
        d_res = torch_mod.full((300 * 32,), 0xAB, dtype=torch_mod.uint8, device="cuda")
        plan = ctx.batch_plan(alg, F_C2, dR, lo, dS, ro, d_res)
        info = plan.info()
        assert info.n_windows == 300 and info.n_items > 0 and info.blocks > 0
        for rep in range(3):
            # eager work and host syncs between replays
            d_res.fill_(0x5A)
            torch_mod.cuda.synchronize()
            plan.launch()
            ctx.sync()
            got = d_res.cpu().numpy().view(sjbmod.WRES_DTYPE)
            assert np.array_equal(got, direct), f"replay {rep}"
        plan.destroy()


def test_autotune_agrees(ctx, torch_mod, sjbmod):
    R, S, lo, ro = batch_case(sjbmod, 128, 0.0, seed=9)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    base = ctx.lazy_join_batch(HJ, F_C2, dR, lo, dS, ro)
    best, tried = ctx.batch_autotune(HJ, F_C2, dR, lo, dS, ro, reps=2)
    # fanout bias {-1,0,1} x chunk {2^12,2^14} x workgroups/CU {1,2,3} x grab {1,4}
    # x match lists on full waves {off, on}
    assert tried == 72 and best.ms > 0
    t = ctx.batch_tune()
    assert (t.fanout_bias, t.chunk_log2, t.blocks_per_cu, t.grab) == \
        (best.fanout_bias, best.chunk_log2, best.blocks_per_cu, best.grab)
    assert np.array_equal(ctx.lazy_join_batch(HJ, F_C2, dR, lo, dS, ro), base)
    ctx.batch_tune(fanout_bias=0, chunk_log2=14, blocks_per_cu=2, grab=1)


def test_tune_validation(ctx):
    with pytest.raises(sjb_error()):
        ctx.batch_tune(chunk_log2=13)
    with pytest.raises(sjb_error()):
        ctx.batch_tune(blocks_per_cu=4)
    with pytest.raises(sjb_error()):
        ctx.batch_tune(expand=2)


@pytest.mark.parametrize("alg,flags", [(HJ, 0), (HJ, F_C2), (SMJ, F_C2)])
def test_expansion_matches(ctx, torch_mod, orc, sjbmod, alg, flags):
    """Batched joins on skewed keys: running long match lists on full waves
    (HJ: every match through the wave's queue; SMJ: cooperative expansion)
    gives the same per-window results as the per-lane loop, and the
    autotuner tries both (72 shapes)."""
    R, S, lo, ro = batch_case(sjbmod, 96, 0.8, seed=23)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    res = {}
    for ex in (0, 1):
        ctx.batch_tune(fanout_bias=0, chunk_log2=12, blocks_per_cu=3, grab=1, expand=ex)
        res[ex] = ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro)
    assert np.array_equal(res[0], res[1])
    for w in (0, 17, 95):
        o, _ = orc.lazy_hj(R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]], 1000)
        want = o.triple() if flags & F_C2 else o.triple()[:2]
        got = (int(res[1][w]["count"]), int(res[1][w]["checksum1"]), int(res[1][w]["checksum2"]))
        assert got[:len(want)] == want
    best, tried = ctx.batch_autotune(alg, flags, dR, lo, dS, ro, reps=1)
    assert tried == 72 and best.expand in (0, 1)
    assert np.array_equal(ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro), res[0])
    ctx.batch_tune(fanout_bias=0, chunk_log2=14, blocks_per_cu=2, grab=1, expand=0)


@pytest.mark.parametrize("alg,flags", [(HJ, F_GROUPED), (SMJ, 0), (SMJ, F_GROUPED)])
def test_expand_ignored_where_it_has_no_variant(ctx, torch_mod, sjbmod, alg, flags):
    """Grouped joins and SMJ without C2 have no full-wave variant: a tune
    with expand = 1 runs the same kernel and gives the same results."""
    R, S, lo, ro = batch_case(sjbmod, 64, 0.8, seed=29)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    res = {}
    for ex in (0, 1):
        ctx.batch_tune(fanout_bias=0, chunk_log2=12, blocks_per_cu=3, grab=1, expand=ex)
        res[ex] = ctx.lazy_join_batch(alg, flags, dR, lo, dS, ro)
    ctx.batch_tune(fanout_bias=0, chunk_log2=14, blocks_per_cu=2, grab=1, expand=0)
    assert np.array_equal(res[0], res[1])


def sjb_error():
    import sjb
    return sjb.SjbError
