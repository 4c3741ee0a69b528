"""GPU tests of the learned partitioner (two-stage RMI with a 1-32-32-1 MLP
root trained on the device).

* The device forward/backward kernel matches the float64 oracle gradient
  (itself pinned by finite differences in test_oracle.py) within a relative
  tolerance of 2e-3 of the gradient's max-norm (float32 accumulation with
  atomics over 4096 samples).
* Joins partitioned by the model are bit-exact (count, C1, C2 / grouped
  count and C1 with peeled heavy keys) against the oracle: the model only
  moves balance, never results.
* What training gains: on the next window of the drifting mixture, the
  trained MLP root gives lower partition imbalance, fewer fix-ups (tuples
  in partitions larger than one LDS tile) and a faster SMJ join than the
  linear-root RMI, which beats radix (range) partitioning on the key bits.
* The model-free baseline (SJB_F_PEEL: heavy keys of the batch's own key
  sample aggregated apart) gives the plain grouped results.
* Drift: a batch plan judges the model on a sample of its own keys and
  refits a stale one to that sample, or falls back to radix ranges with
  SJB_F_NO_REFIT; neither changes the results.
"""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu

HJ, SMJ = 2, 3
F_C2, F_GROUPED, F_LEARNED = 1, 2, 4


def training_windows(sjbmod, n, first, count, seed=40):
    return np.concatenate([sjbmod.generate_mix(n, w, seed=seed + w) for w in range(first, first + count)])


def test_mlp_grad_matches_oracle(ctx, orc):
    rng = np.random.default_rng(11)
    P = rng.normal(0, 0.4, 1153).astype(np.float32)
    x = rng.uniform(0, 1, 4096).astype(np.float32)
    t = np.sort(rng.uniform(0, 1, 4096)).astype(np.float32)
    loss_gpu, g_gpu = ctx.mlp_grad(P, x, t)
    loss_cpu, g_cpu = orc.mlp_grad(P.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    scale = np.abs(g_cpu).max()
    assert scale > 0
    assert np.abs(g_gpu - g_cpu).max() <= 2e-3 * scale
    assert abs(loss_gpu - loss_cpu) <= 1e-5 * max(loss_cpu, 1e-12)


def test_mlp_grad_tiles_and_tail(ctx, orc):
    # n not a multiple of the 128-sample tile: the tail samples must count once
    rng = np.random.default_rng(12)
    P = rng.normal(0, 0.4, 1153).astype(np.float32)
    for n in (1, 127, 129, 1000):
        x = rng.uniform(0, 1, n).astype(np.float32)
        t = rng.uniform(0, 1, n).astype(np.float32)
        _, g_gpu = ctx.mlp_grad(P, x, t)
        _, g_cpu = orc.mlp_grad(P.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        assert np.abs(g_gpu - g_cpu).max() <= 2e-3 * max(np.abs(g_cpu).max(), 1e-30), n


@pytest.fixture(scope="module")
def fitted(sjbmod, torch_mod):
    """A ctx with an MLP-root model fitted on windows 0..3 (2^18 tuples each)."""
    train = training_windows(sjbmod, 1 << 18, 0, 4)
    d_train = to_dev(torch_mod, train)
    c = sjbmod.Context(max_window=1 << 22, distinct_keys=1 << 16)
    rep = c.rmi_fit(d_train, len(train), root=sjbmod.RMI_MLP, steps=400, lr=0.01)
    yield c, rep, d_train, len(train)
    c.close()


def test_fit_report(fitted):
    _, rep, _, _ = fitted
    assert rep.fallback == 0
    assert 1 <= rep.n_heavy <= 32
    assert rep.loss_final < rep.loss_init / 10
    assert rep.imbalance_learned < rep.imbalance_radix
    assert rep.leaves_used > 100


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_learned_grouped_exact(fitted, torch_mod, orc, sjbmod, alg):
    c, _, _, _ = fitted
    n = 1 << 18
    R = sjbmod.generate_mix(n, 4, seed=901)
    S = sjbmod.generate_mix(n, 4, seed=902)
    o, _ = orc.lazy_smj(R, S)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = c.lazy_join(alg, 0, F_GROUPED | F_LEARNED, dR, n, dS, n)
    assert (st.result_count, st.checksum1) == (o.count, o.c1)
    rep = c.partition_report(alg, F_GROUPED | F_LEARNED, dR, n, dS, n)
    # HJ keeps hash partitions and peels the model's heavy keys; SMJ uses CDF ranges
    assert rep.kind == (2 if alg == SMJ else 0) and rep.heavy_left > 0 and rep.heavy_right > 0
    # plain path gives the same result
    st2 = c.lazy_join(alg, 0, F_GROUPED, dR, n, dS, n)
    assert (st2.result_count, st2.checksum1) == (o.count, o.c1)


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_learned_pairwise_exact(fitted, torch_mod, orc, sjbmod, alg):
    # C2 enumerates pairs: heavy keys are partitioned, not peeled
    c, _, _, _ = fitted
    n = 1 << 15
    R = sjbmod.generate_mix(n, 4, seed=911)
    S = sjbmod.generate_mix(n, 4, seed=912)
    o, _ = orc.lazy_hj(R, S, 1 << 16)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = c.lazy_join(alg, 0, F_C2 | F_LEARNED, dR, n, dS, n)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()


def test_learned_batch_equals_single(fitted, torch_mod, sjbmod):
    c, _, _, _ = fitted
    sizes = np.array([1 << 14, 300, 1 << 16, 7, 5000, 1 << 12], dtype=np.uint64)
    lo = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    R = sjbmod.generate_mix(int(lo[-1]), 5, seed=921)
    S = sjbmod.generate_mix(int(lo[-1]), 5, seed=922)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for alg in (HJ, SMJ):
        res = c.lazy_join_batch(alg, F_GROUPED | F_LEARNED, dR, lo, dS, lo)
        for w in range(len(sizes)):
            a, b = int(lo[w]), int(lo[w + 1])
            st = c.lazy_join(alg, 0, F_GROUPED, dR[a * 16:], b - a, dS[a * 16:], b - a)
            assert (int(res[w]["count"]), int(res[w]["checksum1"])) == (st.result_count, st.checksum1)


def join_ms(c, torch_mod, alg, flags, dR, dS, n, reps=5):
    """Median device time (ms) of one grouped window join, and its result."""
    stream = torch_mod.cuda.ExternalStream(c.stream)
    d_res = torch_mod.zeros(32, dtype=torch_mod.uint8, device="cuda")
    c.lazy_join_async(alg, flags, dR, n, dS, n, d_res.data_ptr())
    c.sync()
    ms = []
    for _ in range(reps):
        e0, e1 = torch_mod.cuda.Event(enable_timing=True), torch_mod.cuda.Event(enable_timing=True)
        e0.record(stream)
        c.lazy_join_async(alg, flags, dR, n, dS, n, d_res.data_ptr())
        e1.record(stream)
        c.sync()
        torch_mod.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    r = d_res.cpu().numpy().view(np.uint64)
    return float(np.median(ms)), (int(r[0]), int(r[1]))


def test_trained_network_gains(sjbmod, torch_mod):
    # the next window after the training windows: MLP root vs linear root vs
    # radix, on partition balance and on the SMJ join's device time
    train = training_windows(sjbmod, 1 << 20, 0, 4, seed=70)
    d_train = to_dev(torch_mod, train)
    n = 1 << 22
    R = sjbmod.generate_mix(n, 4, seed=971)
    S = sjbmod.generate_mix(n, 4, seed=972)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    with sjbmod.Context(max_window=n, distinct_keys=1 << 16) as c:
        radix = c.partition_report(SMJ, F_GROUPED, dR, n, dS, n)
        t_radix, res_radix = join_ms(c, torch_mod, SMJ, F_GROUPED, dR, dS, n)
        lin_fit = c.rmi_fit(d_train, len(train), root=sjbmod.RMI_LINEAR)
        lin = c.partition_report(SMJ, F_GROUPED | F_LEARNED, dR, n, dS, n)
        t_lin, res_lin = join_ms(c, torch_mod, SMJ, F_GROUPED | F_LEARNED, dR, dS, n)
        mlp_fit = c.rmi_fit(d_train, len(train), root=sjbmod.RMI_MLP, steps=400, lr=0.01)
        mlp = c.partition_report(SMJ, F_GROUPED | F_LEARNED, dR, n, dS, n)
        t_mlp, res_mlp = join_ms(c, torch_mod, SMJ, F_GROUPED | F_LEARNED, dR, dS, n)
        c.rmi_clear()
        plain = c.partition_report(SMJ, F_GROUPED | F_LEARNED, dR, n, dS, n)
    print("radix", radix.as_dict(), t_radix, "ms\nlinear", lin.as_dict(), lin_fit.as_dict(), t_lin,
          "ms\nmlp", mlp.as_dict(), mlp_fit.as_dict(), t_mlp, "ms", flush=True)
    assert radix.kind == 1 and lin.kind == 2 and mlp.kind == 2 and plain.kind == 1
    assert mlp_fit.loss_final < mlp_fit.loss_init / 10
    assert mlp.imbalance < lin.imbalance < radix.imbalance
    assert mlp.fixup_frac < lin.fixup_frac < radix.fixup_frac
    # the same results, and what the models buy in join time
    assert res_radix == res_lin == res_mlp
    assert t_mlp < t_radix / 1.5 and t_lin < t_radix / 1.2, (t_radix, t_lin, t_mlp)


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_sampled_peeling_exact(ctx, torch_mod, sjbmod, alg):
    # SJB_F_PEEL (no model): heavy keys of the batch's own sample are
    # aggregated apart; per-window results equal the plain grouped join
    n, nw = 1 << 16, 6
    lo = np.arange(nw + 1, dtype=np.uint64) * np.uint64(n)
    R = np.concatenate([sjbmod.generate_mix(n, w, seed=981 + w) for w in range(nw)])
    S = np.concatenate([sjbmod.generate_mix(n, w, seed=991 + w) for w in range(nw)])
    R["key"][::1001] = 0xFFFFFFFF
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    got = {}
    for f in (F_GROUPED, F_GROUPED | sjbmod.F_PEEL):
        d_res = torch_mod.zeros(nw * 32, dtype=torch_mod.uint8, device="cuda")
        plan = ctx.batch_plan(alg, f, dR, lo, dS, lo, d_res)
        plan.launch()
        ctx.sync()
        info = plan.info()
        assert info.n_heavy == (0 if f == F_GROUPED else info.n_heavy) and (f == F_GROUPED or info.n_heavy > 0)
        assert info.part_kind == (1 if alg == SMJ else 0)
        got[f] = d_res.cpu().numpy().view(sjbmod.WRES_DTYPE)[["count", "checksum1"]].copy()
        plan.destroy()
    assert np.array_equal(got[F_GROUPED], got[F_GROUPED | sjbmod.F_PEEL])


def test_drift_judge_refits_or_falls_back(fitted, torch_mod, sjbmod):
    # The model was fitted on windows 0..3; 640 windows later the mixture
    # means have moved by 10% of the key range and the learned CDF is
    # stale: the batch plan judges it on a sample of its own keys and
    # refits a private copy to that sample (learned CDF kept), or with
    # SJB_F_NO_REFIT uses radix ranges (heavy keys stay peeled either way);
    # results do not change.
    c, _, _, _ = fitted
    n, nw = 1 << 16, 4
    lo = np.arange(nw + 1, dtype=np.uint64) * np.uint64(n)
    F_NO_REFIT = sjbmod.F_NO_REFIT
    for first, stale in ((4, False), (640, True)):
        R = np.concatenate([sjbmod.generate_mix(n, w, seed=931 + w) for w in range(first, first + nw)])
        S = np.concatenate([sjbmod.generate_mix(n, w, seed=951 + w) for w in range(first, first + nw)])
        dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
        got = {}
        for f in (F_GROUPED | F_LEARNED, F_GROUPED | F_LEARNED | F_NO_REFIT, F_GROUPED):
            d_res = torch_mod.zeros(nw * 32, dtype=torch_mod.uint8, device="cuda")
            plan = c.batch_plan(SMJ, f, dR, lo, dS, lo, d_res)
            plan.launch()
            c.sync()
            info = plan.info()
            if not f & F_LEARNED:
                assert (info.drift_fallback, info.drift_refit, info.part_kind) == (0, 0, 1)
            elif not stale:
                assert (info.drift_fallback, info.drift_refit, info.part_kind) == (0, 0, 2), first
            elif f & F_NO_REFIT:
                assert (info.drift_fallback, info.drift_refit, info.part_kind) == (1, 0, 1)
            else:
                assert (info.drift_fallback, info.drift_refit, info.part_kind) == (0, 1, 2)
            got[f] = d_res.cpu().numpy().view(sjbmod.WRES_DTYPE)[["count", "checksum1"]].copy()
            plan.destroy()
        assert np.array_equal(got[F_GROUPED | F_LEARNED], got[F_GROUPED])
        assert np.array_equal(got[F_GROUPED | F_LEARNED | F_NO_REFIT], got[F_GROUPED])


def test_fit_after_sampled_peeling(sjbmod, torch_mod):
    # a batch plan with SJB_F_PEEL samples keys before any fit; a later fit
    # on the same context must still get every fit buffer
    n, nw = 1 << 16, 2
    lo = np.arange(nw + 1, dtype=np.uint64) * np.uint64(n)
    R = np.concatenate([sjbmod.generate_mix(n, w, seed=961 + w) for w in range(nw)])
    S = np.concatenate([sjbmod.generate_mix(n, w, seed=971 + w) for w in range(nw)])
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    with sjbmod.Context(max_window=n) as c:
        d_res = torch_mod.zeros(nw * 32, dtype=torch_mod.uint8, device="cuda")
        plan = c.batch_plan(HJ, F_GROUPED | sjbmod.F_PEEL, dR, lo, dS, lo, d_res)
        plan.launch()
        c.sync()
        assert plan.info().n_heavy > 0
        plan.destroy()
        rep = c.rmi_fit(dR, len(R), root=sjbmod.RMI_MLP, steps=50, lr=0.01)
        assert rep.fallback == 0 and rep.loss_final < rep.loss_init
        st = c.lazy_join(SMJ, 0, F_GROUPED | F_LEARNED, dR, n, dS, n)
        st2 = c.lazy_join(SMJ, 0, F_GROUPED, dR, n, dS, n)
        assert (st.result_count, st.checksum1) == (st2.result_count, st2.checksum1)
