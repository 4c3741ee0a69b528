"""GPU parity: the HIP operators (through the C ABI) against the CPU oracle.

Bar: bit-exact count, C1 and C2 (integer arithmetic mod 2^64) and identical
sorted pair lists wherever pairs are produced.
"""
import numpy as np
import pytest

from conftest import to_dev
from helpers import make, np_join, np_pair_rows, sort_rows, zipf_keys

pytestmark = pytest.mark.gpu

HJ, SMJ, NLJ = 2, 3, 1
NO, CK, ES, BM, AT = 0, 1, 2, 3, 4
F_C2, F_GROUPED = 1, 2

SIZES = [(1, 1), (100, 37), (5000, 4096), (16389, 20011), (70000, 65536)]


def run_count(ctx, torch, alg, R, S, flags=F_C2):
    dR, dS = to_dev(torch, R), to_dev(torch, S)
    st = ctx.lazy_join(alg, NO, flags, dR, len(R), dS, len(S))
    return st


@pytest.mark.parametrize("alg", [HJ, SMJ, NLJ])
@pytest.mark.parametrize("key_range", [16, 1000, 1 << 16, 1 << 24])
@pytest.mark.parametrize("size", SIZES)
def test_count_checksums(ctx, torch_mod, orc, alg, key_range, size):
    nL, nR = size
    print(f"case alg={alg} range={key_range} size={size}", flush=True)
    R, S = make(nL, key_range, 1000 + nL), make(nR, key_range, 2000 + nR)
    ref = np_join(R, S) if nL * nR < 1 << 26 or key_range > 1 << 12 else None
    o, _ = orc.lazy_hj(R, S, 1 << 16)
    if ref is not None:
        assert o.triple() == ref
    st = run_count(ctx, torch_mod, alg, R, S)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    assert st.c2_valid == 1


@pytest.mark.parametrize("alg", [HJ, SMJ])
@pytest.mark.parametrize("theta", [0.8, 1.2])
def test_skewed_keys(ctx, torch_mod, orc, alg, theta):
    n = 1 << 16
    print(f"case skew alg={alg} theta={theta}", flush=True)
    R = make(n, 1, 11, keys=zipf_keys(n, theta, 1 << 14, 1))
    S = make(n, 1, 12, keys=zipf_keys(n, theta, 1 << 14, 2))
    o, _ = orc.lazy_smj(R, S)
    for flags in (0, F_GROUPED):
        st = run_count(ctx, torch_mod, alg, R, S, flags)
        assert (st.result_count, st.checksum1) == (o.count, o.c1)
        assert st.c2_valid == 0


def test_grouped_equals_pairwise(ctx, torch_mod):
    R, S = make(1 << 18, 1 << 12, 5), make(1 << 18, 1 << 12, 6)
    a = run_count(ctx, torch_mod, SMJ, R, S, F_GROUPED)
    b = run_count(ctx, torch_mod, SMJ, R, S, F_C2)
    c = run_count(ctx, torch_mod, HJ, R, S, F_C2)
    d = run_count(ctx, torch_mod, HJ, R, S, F_GROUPED)
    assert (a.result_count, a.checksum1) == (b.result_count, b.checksum1)
    assert (d.result_count, d.checksum1) == (b.result_count, b.checksum1)
    assert (b.result_count, b.checksum1, b.checksum2) == (c.result_count, c.checksum1, c.checksum2)


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_grouped_key_ffffffff(ctx, torch_mod, orc, alg):
    # the aggregation table's empty marker is a legal key
    R, S = make(5000, 40, 141), make(4000, 40, 142)
    R["key"][::3] = 0xFFFFFFFF
    S["key"][::5] = 0xFFFFFFFF
    o, _ = orc.lazy_smj(R, S)
    st = run_count(ctx, torch_mod, alg, R, S, F_GROUPED)
    assert (st.result_count, st.checksum1) == (o.count, o.c1)


@pytest.mark.parametrize("alg,method", [(HJ, CK), (HJ, ES), (HJ, AT), (SMJ, CK), (SMJ, ES),
                                        (SMJ, AT), (NLJ, CK), (NLJ, ES), (NLJ, AT), (NLJ, BM)])
def test_materialised_pairs(ctx, torch_mod, orc, alg, method):
    print(f"case pairs alg={alg} method={method}", flush=True)
    R, S = make(3000, 500, 71), make(2500, 500, 72)
    ref_rows = sort_rows(np_pair_rows(R, S))
    o, _ = orc.lazy_nlj(R, S)
    cap = len(ref_rows) + 10
    out = torch_mod.zeros(cap * 32, dtype=torch_mod.uint8, device="cuda")
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = ctx.lazy_join(alg, method, 0, dR, len(R), dS, len(S), out, cap)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    assert st.pairs_written == len(ref_rows)
    rows = out.cpu().numpy().view(ref_rows.dtype)[: st.pairs_written]
    assert np.array_equal(sort_rows(rows), ref_rows)


@pytest.mark.parametrize("alg", [HJ, SMJ])
@pytest.mark.parametrize("size", [((1 << 19) - 5, (1 << 19) + 5), ((1 << 19) + 3, 1 << 19),
                                  ((1 << 16) - 1, (1 << 16) + 1), ((1 << 16) + 1, 1 << 16)])
def test_small_window_boundary(ctx, torch_mod, orc, alg, size):
    """Partition chunk sizes switch with the window's total size: exactly
    2^20 tuples (both sides) take 2048-tuple chunks and three more the
    16384-tuple chunks of large windows; exactly 2^17 take 1024-tuple chunks
    and one more 2048-tuple chunks.  All give the oracle's triple."""
    nL, nR = size
    print(f"case boundary alg={alg} size={size}", flush=True)
    R, S = make(nL, 1 << 22, 500 + nL), make(nR, 1 << 22, 600 + nR)
    o, _ = orc.lazy_hj(R, S, 1 << 16)
    st = run_count(ctx, torch_mod, alg, R, S)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    st = run_count(ctx, torch_mod, alg, R, S, F_GROUPED)
    assert (st.result_count, st.checksum1) == (o.count, o.c1)


@pytest.mark.parametrize("alg", [HJ, SMJ, NLJ])
def test_small_context_windows(sjbmod, torch_mod, orc, alg):
    """A context made for 2^16-tuple windows (its offsets buffer sized for
    the 1024- and 2048-tuple chunks of small windows): full, ragged and tiny
    windows, plain and asynchronous, give the oracle's triple."""
    W = 1 << 16
    with sjbmod.Context(device=0, max_window=W, distinct_keys=1 << 16) as c:
        for nL, nR in [(W, W), (W - 1, 1000), (3, W)]:
            R, S = make(nL, 1 << 16, 700 + nL), make(nR, 1 << 16, 800 + nR)
            o, _ = orc.lazy_hj(R, S, 1 << 16)
            st = run_count(c, torch_mod, alg, R, S)
            assert (st.result_count, st.checksum1, st.checksum2) == o.triple(), (nL, nR)
            dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
            d = torch_mod.zeros(32, dtype=torch_mod.uint8, device="cuda")
            c.lazy_join_async(alg, F_C2, dR.data_ptr(), nL, dS.data_ptr(), nR, d.data_ptr())
            c.sync()
            r = d.cpu().numpy().view(sjbmod.WRES_DTYPE)[0]
            assert (int(r["count"]), int(r["checksum1"]), int(r["checksum2"])) == o.triple(), (nL, nR)


@pytest.mark.parametrize("key_range", [16, 1000, 1 << 24])
@pytest.mark.parametrize("size", [(1, 1), (100, 37), (1023, 1025), (5000, 4096), (16389, 20011)])
def test_bitmap_pairs(ctx, torch_mod, orc, key_range, size):
    """NLJ Bitmap at ragged sizes (partial right rows and left tiles, one
    and several tiles per workgroup) and dense to sparse matches: every
    (left, 64-key word) mask is found once, so rows and checksums equal the
    pair enumeration's."""
    nL, nR = size
    if key_range == 16 and nL * nR > 1 << 23:
        nL, nR = 1501, 1499  # keep the dense case's pair list small
    print(f"case bitmap range={key_range} size={(nL, nR)}", flush=True)
    R, S = make(nL, key_range, 3000 + nL), make(nR, key_range, 4000 + nR)
    ref_rows = sort_rows(np_pair_rows(R, S))
    o, _ = orc.lazy_hj(R, S, 1 << 16)
    cap = len(ref_rows) + 10
    out = torch_mod.zeros(cap * 32, dtype=torch_mod.uint8, device="cuda")
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = ctx.lazy_join(NLJ, BM, 0, dR, nL, dS, nR, out, cap)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    assert st.pairs_written == len(ref_rows)
    rows = out.cpu().numpy().view(ref_rows.dtype)[: st.pairs_written]
    assert np.array_equal(sort_rows(rows), ref_rows)


def test_bitmap_window16(ctx, torch_mod, orc, sjbmod):
    """NLJ Bitmap on a full 2^16 x 2^16 window (keys uniform in [0, 2^24),
    ~256 pairs): rows and checksums against the pair enumeration and the
    oracle, and the count equals the count kernel's."""
    W = 1 << 16
    R = sjbmod.generate(W, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=777)
    S = sjbmod.generate(W, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=778)
    ref_rows = sort_rows(np_pair_rows(R, S))
    o, _ = orc.lazy_hj(R, S, 1 << 16)
    cap = len(ref_rows) + 10
    out = torch_mod.zeros(cap * 32, dtype=torch_mod.uint8, device="cuda")
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = ctx.lazy_join(NLJ, BM, 0, dR, W, dS, W, out, cap)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    rows = out.cpu().numpy().view(ref_rows.dtype)[: st.pairs_written]
    assert np.array_equal(sort_rows(rows), ref_rows)
    assert run_count(ctx, torch_mod, NLJ, R, S).result_count == o.count


@pytest.mark.parametrize("spread", [1 << 12, (1 << 20) - 1, 1 << 21, 1 << 32])
@pytest.mark.parametrize("size", [(1000, 900), (3000, 2500), (5000, 4096)])
def test_smj_sort_key_spans(ctx, torch_mod, orc, spread, size):
    """SMJ build tiles whose keys span less / more than 2^20 (the sort runs
    on packed 32-bit values below that span, on the 64-bit records above
    it); keys from a pool that includes 0 and spread - 1 (0xFFFFFFFF for the
    full range), so every tile has matches."""
    nL, nR = size
    print(f"case smj spans spread={spread} size={size}", flush=True)
    rng = np.random.default_rng(nL + spread % 977)
    pool = rng.integers(0, spread, 300, dtype=np.uint64)
    pool[0], pool[1] = 0, spread - 1
    pool = pool.astype(np.uint32)
    R = make(nL, 1, 31 + nL, keys=pool[rng.integers(0, len(pool), nL)])
    S = make(nR, 1, 37 + nR, keys=pool[rng.integers(0, len(pool), nR)])
    o, _ = orc.lazy_hj(R, S, 1 << 12)
    st = run_count(ctx, torch_mod, SMJ, R, S)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    if size == (3000, 2500):
        ref_rows = sort_rows(np_pair_rows(R, S))
        cap = len(ref_rows) + 10
        out = torch_mod.zeros(cap * 32, dtype=torch_mod.uint8, device="cuda")
        st = ctx.lazy_join(SMJ, AT, 0, to_dev(torch_mod, R), nL, to_dev(torch_mod, S), nR, out, cap)
        assert st.pairs_written == len(ref_rows)
        rows = out.cpu().numpy().view(ref_rows.dtype)[: st.pairs_written]
        assert np.array_equal(sort_rows(rows), ref_rows)


def test_smj_range_bounds_from_sample(ctx, torch_mod, orc):
    """Range partitioning takes its key bounds from a sample of at most 2^16
    keys; keys outside them are clamped to the end partitions.  Matching
    outlier keys (far below and far above the bulk) are put only at
    positions the sample skips, and must still meet."""
    nL, nR = 150000, 130000
    n, ns = nL + nR, 1 << 16
    sampled = np.zeros(n, dtype=bool)
    sampled[(np.arange(ns, dtype=np.uint64) * n) // ns] = True
    free = np.flatnonzero(~sampled)
    rng = np.random.default_rng(5)
    R = make(nL, 1, 61, keys=(1000 + rng.integers(0, 1 << 16, nL)).astype(np.uint32))
    S = make(nR, 1, 62, keys=(1000 + rng.integers(0, 1 << 16, nR)).astype(np.uint32))
    fl, fr = free[free < nL], free[free >= nL] - nL
    for key, cl, cr in ((5, 3, 2), (0xFFFFFFF0, 4, 3), (0xFFFFFFFF, 2, 2)):
        R["key"][rng.choice(fl, cl, replace=False)] = key
        S["key"][rng.choice(fr, cr, replace=False)] = key
    o, _ = orc.lazy_hj(R, S, 1 << 12)
    st = run_count(ctx, torch_mod, SMJ, R, S)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    assert o.count >= 3 * 2 + 4 * 3 + 2 * 2


@pytest.mark.parametrize("alg", [HJ, SMJ, NLJ])
def test_capacity_errors(ctx, torch_mod, alg):
    R, S = make(2000, 50, 81), make(2000, 50, 82)
    n_ref = np_join(R, S)[0]
    out = torch_mod.zeros(100 * 32, dtype=torch_mod.uint8, device="cuda")
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for method in (CK, ES):
        rc, st = ctx.lazy_join(alg, method, 0, dR, len(R), dS, len(S), out, 100,
                               raise_on_error=False)
        assert rc == -2 and st.result_count == n_ref
    # Atomic wraps like the reference's sink ring: no error, count exact
    rc, st = ctx.lazy_join(alg, AT, 0, dR, len(R), dS, len(S), out, 100, raise_on_error=False)
    assert rc == 0 and st.result_count == n_ref and st.pairs_written == 100
    # window larger than the context
    rc, st = ctx.lazy_join(alg, NO, 0, dR, (1 << 22) + 1, dS, 10, raise_on_error=False)
    assert rc == -2
    rc, st = ctx.lazy_join(SMJ if alg == HJ else HJ, BM, 0, dR, 10, dS, 10, out, 100,
                           raise_on_error=False)
    assert rc == -5


@pytest.mark.parametrize("logF", [1, 4, 9, 12])
def test_fanout_invariance(sjbmod, torch_mod, orc, logF):
    R, S = make(50000, 1 << 14, 91), make(60000, 1 << 14, 92)
    o, _ = orc.lazy_hj(R, S, 1 << 14)
    with sjbmod.Context(max_window=1 << 17, fanout_log2=logF) as c:
        for alg in (HJ, SMJ):
            st = run_count(c, torch_mod, alg, R, S)
            assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
            assert st.fanout_log2 == logF


def test_lazy_execute_ring_slots(ctx, orc):
    B = 512
    ringL, ringR = make(64 * B, 300, 101), make(64 * B, 300, 102)
    li = np.array([5, 6, 7, 40, 2, 3], dtype=np.uint64) * B  # non-adjacent slots
    ri = np.array([63, 0, 1, 10, 11, 12], dtype=np.uint64) * B
    L = np.concatenate([ringL[i:i + B] for i in li])
    R = np.concatenate([ringR[i:i + B] for i in ri])
    o, _ = orc.lazy_hj(L, R, 1000)
    for alg in (HJ, SMJ, NLJ):
        rc, st = ctx.lazy_execute(alg, NO, F_C2, ringL, li, ringR, ri, B)
        assert rc == 0
        assert (st.result_count, st.checksum1, st.checksum2) == o.triple()


def test_batch_equals_single(ctx, torch_mod, orc):
    rng = np.random.default_rng(5)
    sizes_l = rng.integers(0, 40000, 12)
    sizes_r = rng.integers(0, 40000, 12)
    sizes_l[3] = 0
    R = make(int(sizes_l.sum()), 5000, 111)
    S = make(int(sizes_r.sum()), 5000, 112)
    lo = np.concatenate([[0], np.cumsum(sizes_l)]).astype(np.uint64)
    so = np.concatenate([[0], np.cumsum(sizes_r)]).astype(np.uint64)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    for alg in (HJ, SMJ):
        res = ctx.lazy_join_batch(alg, F_C2, dR, lo, dS, so)
        for w in range(12):
            o, _ = orc.lazy_hj(R[lo[w]:lo[w + 1]], S[so[w]:so[w + 1]], 1000)
            assert (int(res[w]["count"]), int(res[w]["checksum1"]), int(res[w]["checksum2"])) == o.triple()


@pytest.mark.parametrize("alg", [HJ, NLJ])
def test_eager_vs_oracle(sjbmod, orc, alg):
    B, nb = 1000, 10
    W = B * nb
    R, S = make(W, 400, 121), make(W, 400, 122)
    e = orc.Eager(1000, W)
    with sjbmod.Context(max_window=W) as c:
        for rnd in range(2):  # two windows: clear between them
            for b in range(nb):
                for is_left, src in ((1, R), (0, S)):
                    batch = src[b * B:(b + 1) * B]
                    rc, st = c.eager_execute(alg, NO, F_C2, batch, is_left)
                    assert rc == 0
                    o = e.execute(batch, is_left)
                    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
            c.eager_clear()
            e.clear()
        # capacity: a full side rejects one more batch
        for b in range(nb):
            c.eager_execute(alg, NO, 0, R[b * B:(b + 1) * B], 1)
        rc, _ = c.eager_execute(alg, NO, 0, R[:B], 1)
        assert rc == -2


def test_eager_atomic_pairs(sjbmod, orc, torch_mod):
    B = 700
    R, S = make(4 * B, 200, 131), make(4 * B, 200, 132)
    with sjbmod.Context(max_window=4 * B) as c:
        rows = []
        for b in range(4):
            for is_left, src in ((1, R), (0, S)):
                out = np.zeros(B * 4 * B, dtype=sjbmod.RESULT_DTYPE)
                rc, st = c.eager_execute(HJ, AT, 0, src[b * B:(b + 1) * B], is_left, out, len(out))
                assert rc == 0
                rows.append(out[: st.pairs_written])
        got = sort_rows(np.concatenate(rows))
        assert np.array_equal(got, sort_rows(np_pair_rows(R, S)))


def test_full_size_window(ctx, torch_mod, orc, sjbmod):
    # 2^22 tuples per side, keys uniform in [0, 2^24): HJ == SMJ == oracle.
    n = 1 << 22
    R = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=123)
    S = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=456)
    o, _ = orc.lazy_hj(R, S, 1 << 20)
    for alg in (HJ, SMJ):
        st = run_count(ctx, torch_mod, alg, R, S)
        assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    assert 0.9 * 2**20 < o.count < 1.1 * 2**20


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_resident_join_walk(ctx, torch_mod, orc, sjbmod, alg):
    # 2^21 tuples per side: HJ 2^10 (grouped 2^12) and SMJ 2^11 partitions,
    # more than the resident count join's 768 workgroups, so each walks
    # several partitions of its XCD; count + C1 only, count + C1 + C2 and
    # grouped.
    n = 1 << 21
    print(f"case resident walk alg={alg} n={n}", flush=True)
    R = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 23, 0.0, seed=41)
    S = sjbmod.generate(n, sjbmod.DIST_UNIFORM, 1 << 23, 0.0, seed=42)
    o, _ = orc.lazy_hj(R, S, 1 << 20)
    st = run_count(ctx, torch_mod, alg, R, S, flags=0)
    assert (st.result_count, st.checksum1) == o.triple()[:2]
    assert st.fanout_log2 >= 10
    st = run_count(ctx, torch_mod, alg, R, S)
    assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
    st = run_count(ctx, torch_mod, alg, R, S, flags=F_GROUPED)
    assert (st.result_count, st.checksum1) == o.triple()[:2]


def test_resident_join_skew(ctx, torch_mod, orc, sjbmod):
    # Zipf(0.5) keys over [0, 2^22), 2^21 tuples per side: partitions of
    # uneven sizes (the hottest key has ~500 copies per side, ~2 matches per
    # tuple) walked by the resident join.
    n = 1 << 21
    print(f"case resident skew n={n}", flush=True)
    R = sjbmod.generate(n, sjbmod.DIST_ZIPF, 1 << 22, 0.5, seed=43)
    S = sjbmod.generate(n, sjbmod.DIST_ZIPF, 1 << 22, 0.5, seed=44)
    o, _ = orc.lazy_hj(R, S, 1 << 20)
    for alg in (HJ, SMJ):
        st = run_count(ctx, torch_mod, alg, R, S)
        assert (st.result_count, st.checksum1, st.checksum2) == o.triple(), alg
        st = run_count(ctx, torch_mod, alg, R, S, flags=F_GROUPED)
        assert (st.result_count, st.checksum1) == o.triple()[:2], alg
    assert o.count > n


def test_window_above_2_22(torch_mod, orc, sjbmod):
    # 2^23 + 12345 / 2^23 - 777 tuples per side (keys uniform in [0, 2^25)):
    # HJ at fanout 2^12 with half-full tiles and SMJ at 2^12, 513 / 512
    # chunks per side, so the join builds its run table in several passes
    # (the general view path) at full scale.
    nL, nR = (1 << 23) + 12345, (1 << 23) - 777
    print(f"case window above 2^22 nL={nL} nR={nR}", flush=True)
    R = sjbmod.generate(nL, sjbmod.DIST_UNIFORM, 1 << 25, 0.0, seed=31)
    S = sjbmod.generate(nR, sjbmod.DIST_UNIFORM, 1 << 25, 0.0, seed=32)
    o, _ = orc.lazy_hj(R, S, 1 << 20)
    with sjbmod.Context(device=0, max_window=1 << 24, distinct_keys=1 << 25) as c:
        for alg in (HJ, SMJ):
            st = run_count(c, torch_mod, alg, R, S)
            assert (st.result_count, st.checksum1, st.checksum2) == o.triple(), alg
            assert st.fanout_log2 == 12, alg
    assert 0.9 * 2**21 < o.count < 1.1 * 2**21


def test_ragged_multichunk_windows(ctx, torch_mod, orc, sjbmod):
    # more 16384-tuple chunks than partition workgroups (one per CU), both
    # sides ending in a partial chunk and of different lengths: the
    # persistent partition pass walks several chunks per workgroup, with
    # the next chunk's loads in flight
    nL, nR = 3 * (1 << 20) + 12345, (1 << 22) - 777
    R = sjbmod.generate(nL, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=321)
    S = sjbmod.generate(nR, sjbmod.DIST_UNIFORM, 1 << 24, 0.0, seed=654)
    o, _ = orc.lazy_hj(R, S, 1 << 20)
    for alg in (HJ, SMJ):
        st = run_count(ctx, torch_mod, alg, R, S)
        assert (st.result_count, st.checksum1, st.checksum2) == o.triple(), alg
        st = run_count(ctx, torch_mod, alg, R, S, flags=F_GROUPED)
        assert (st.result_count, st.checksum1) == (o.count, o.c1), alg


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_overlap_async_equals_plain(sjbmod, torch_mod, orc, alg):
    # back-to-back async windows with the partition pass of window w+1 on a
    # second stream (SJB_F_OVERLAP) give the same per-window results; a
    # plain call in between and the timing ring keep working
    W, nw = 1 << 16, 9
    R = sjbmod.generate(W * nw, sjbmod.DIST_UNIFORM, 1 << 15, 0.0, seed=61)
    S = sjbmod.generate(W * nw, sjbmod.DIST_UNIFORM, 1 << 15, 0.0, seed=62)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    with sjbmod.Context(max_window=W) as c:
        c.timing_ring(2 * nw)
        outs = []
        for flags in (F_C2, F_C2 | sjbmod.F_OVERLAP | sjbmod.F_TIMING):
            d = torch_mod.zeros(nw * 32, dtype=torch_mod.uint8, device="cuda")
            for w in range(nw):
                c.lazy_join_async(alg, flags, dR.data_ptr() + w * W * 16, W, dS.data_ptr() + w * W * 16, W,
                                  d.data_ptr() + w * 32)
                if w == 4:
                    st = c.lazy_join(alg, 0, F_C2, dR, W, dS, W)  # plain call on buffer set 0
            c.sync()
            outs.append(d.cpu().numpy().view(sjbmod.WRES_DTYPE).copy())
        assert np.array_equal(outs[0], outs[1])
        part, join = c.timing_read(nw)
        assert len(part) == nw and np.all(part > 0) and np.all(join > 0)
    for w in (0, nw - 1):
        o, _ = orc.lazy_hj(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W], 1 << 12)
        assert (int(outs[1][w]["count"]), int(outs[1][w]["checksum1"]), int(outs[1][w]["checksum2"])) == o.triple()


def test_pinned_ring_slots_and_output(ctx, sjbmod, orc):
    """Ring slots and pair rows in page-locked memory from sjb_host_alloc
    (asynchronous copies in, device writes straight into the host rows)."""
    import ctypes
    L = sjbmod.lib()
    B, nslots = 1024, 16
    ringL, ringR, outp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert L.sjb_host_alloc(nslots * B * 16, ctypes.byref(ringL)) == 0
    assert L.sjb_host_alloc(nslots * B * 16, ctypes.byref(ringR)) == 0
    cap = 1 << 16
    assert L.sjb_host_alloc(cap * 32, ctypes.byref(outp)) == 0
    try:
        rl = np.ctypeslib.as_array((ctypes.c_uint8 * (nslots * B * 16)).from_address(ringL.value)).view(sjbmod.TUPLE_DTYPE)
        rr = np.ctypeslib.as_array((ctypes.c_uint8 * (nslots * B * 16)).from_address(ringR.value)).view(sjbmod.TUPLE_DTYPE)
        rl[:] = make(nslots * B, 700, 201)
        rr[:] = make(nslots * B, 700, 202)
        li = np.array([3, 4, 5, 9, 0], dtype=np.uint64) * B
        ri = np.array([15, 1, 2, 7, 8], dtype=np.uint64) * B
        Lw = np.concatenate([rl[i:i + B] for i in li])
        Rw = np.concatenate([rr[i:i + B] for i in ri])
        o, _ = orc.lazy_hj(Lw, Rw, 1000)
        for alg in (HJ, SMJ, NLJ):
            rc, st = ctx.lazy_execute(alg, NO, F_C2, ringL.value, li, ringR.value, ri, B)
            assert rc == 0 and (st.result_count, st.checksum1, st.checksum2) == o.triple()
        st = ctx.lazy_join(HJ, AT, 0, to_dev_np(Lw), len(Lw), to_dev_np(Rw), len(Rw), outp.value, cap)
        rows = np.ctypeslib.as_array((ctypes.c_uint8 * (cap * 32)).from_address(outp.value))
        rows = rows.view(np.dtype([("lkey", "<u4"), ("rkey", "<u4"), ("lval", "<u4"), ("rval", "<u4"),
                                   ("lts", "<u8"), ("rts", "<u8")]))[: st.pairs_written]
        assert st.pairs_written == o.count
        assert np.array_equal(sort_rows(rows.copy()), sort_rows(np_pair_rows(Lw, Rw)))
    finally:
        for p in (ringL, ringR, outp):
            assert L.sjb_host_free(p) == 0


def to_dev_np(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to("cuda")


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_async_result_slots_start_dirty(ctx, torch_mod, orc, sjbmod, alg):
    """The count paths sum into the caller's result slot (the partition
    pass zeroes it, each join workgroup adds its share): slots full of stale
    bytes, with several windows of different sizes in flight on the stream,
    still get exact count / C1 / C2."""
    sizes = [(1 << 16, 1 << 16), (70000, 65536), (5000, 4096), (1 << 20, 1 << 20)]
    wins = [(make(nl, 1 << 16, 300 + i), make(nr, 1 << 16, 400 + i)) for i, (nl, nr) in enumerate(sizes)]
    refs = [orc.lazy_hj(R, S, 1 << 16)[0] for R, S in wins]
    dev = [(to_dev(torch_mod, R), to_dev(torch_mod, S)) for R, S in wins]
    for flags in (F_C2, F_GROUPED, 0):
        print(f"case dirty slots alg={alg} flags={flags}", flush=True)
        d_res = torch_mod.full((32 * len(wins),), 0xAB, dtype=torch_mod.uint8, device="cuda")
        for i, ((dR, dS), (R, S)) in enumerate(zip(dev, wins)):
            ctx.lazy_join_async(alg, flags, dR, len(R), dS, len(S), d_res.data_ptr() + 32 * i)
        ctx.sync()
        got = d_res.cpu().numpy().view(sjbmod.WRES_DTYPE)
        for i, o in enumerate(refs):
            assert (int(got[i]["count"]), int(got[i]["checksum1"])) == (o.count, o.c1), (flags, i)
            if flags & F_C2:
                assert int(got[i]["checksum2"]) == o.c2, (flags, i)
