"""The HIP operators against the committed golden fixtures directly (no
oracle call at run time): the small windows' count / C1 / C2 for every
algorithm and their materialised pair lists, and cfg1's 16 windows through
the asynchronous per-window path."""
import numpy as np
import pytest

from conftest import to_dev
from golden_io import WINDOWS, load_cfg1, load_window, rows_sorted

pytestmark = pytest.mark.gpu

NLJ, HJ, SMJ = 1, 2, 3
NO, CK = 0, 1
F_C2 = 1


@pytest.mark.parametrize("name", WINDOWS)
@pytest.mark.parametrize("alg", [HJ, SMJ, NLJ])
def test_window_fixture(ctx, torch_mod, name, alg):
    print(f"case golden {name} alg={alg}", flush=True)
    R, S, triple, pairs = load_window(name)
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    st = ctx.lazy_join(alg, NO, F_C2, dR, len(R), dS, len(S))
    assert (st.result_count, st.checksum1, st.checksum2) == triple
    cap = len(pairs) + 8
    out = torch_mod.zeros(cap * 32, dtype=torch_mod.uint8, device="cuda")
    st = ctx.lazy_join(alg, CK, 0, dR, len(R), dS, len(S), out, cap)
    assert st.pairs_written == len(pairs)
    rows = out.cpu().numpy().view(np.dtype([("lkey", "<u4"), ("rkey", "<u4"), ("lval", "<u4"),
                                            ("rval", "<u4"), ("lts", "<u8"), ("rts", "<u8")]))
    rows = rows[: st.pairs_written]
    assert np.array_equal(rows_sorted(rows["lkey"], rows["lval"], rows["rval"]), pairs)


@pytest.mark.parametrize("alg", [HJ, SMJ])
def test_cfg1_fixture(sjbmod, torch_mod, alg):
    doc = load_cfg1()
    g, W = doc["generator"], doc["window"]
    print(f"case golden cfg1 alg={alg}", flush=True)
    R = sjbmod.generate(g["n"], sjbmod.DIST_UNIFORM, g["key_range"], 0.0, seed=g["seed_left"])
    S = sjbmod.generate(g["n"], sjbmod.DIST_UNIFORM, g["key_range"], 0.0, seed=g["seed_right"])
    dR, dS = to_dev(torch_mod, R), to_dev(torch_mod, S)
    nw = len(doc["windows"])
    d_res = torch_mod.zeros(32 * nw, dtype=torch_mod.uint8, device="cuda")
    with sjbmod.Context(device=0, max_window=W, distinct_keys=g["key_range"]) as c:
        for w in range(nw):
            c.lazy_join_async(alg, F_C2, dR.data_ptr() + w * W * 16, W, dS.data_ptr() + w * W * 16, W,
                              d_res.data_ptr() + 32 * w)
        c.sync()
    got = d_res.cpu().numpy().view(sjbmod.WRES_DTYPE)
    for w, ref in enumerate(doc["windows"]):
        assert (int(got[w]["count"]), int(got[w]["checksum1"]), int(got[w]["checksum2"])) == \
            (ref["count"], ref["c1"], ref["c2"]), w
