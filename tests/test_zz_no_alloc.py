"""GPU test of sjb_lazy_join_noalloc, the allocation-free window join.

Its per-window results (count, C1, and C2 or the grouped count + C1) must be
bit-identical to sjb_lazy_join_async's on the same inputs, for HJ and SMJ,
over two calls in a row on one stream — the first at the maximum size of
2^20 tuples per side, the second smaller, so the second call runs on what
the first left in the shared scratch array.  One case is also checked
against the oracle.  Sizes above the maximum and unsupported requests are
rejected before any GPU work.
"""
import ctypes

import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu

HJ, SMJ, NLJ = 2, 3, 1
F_C2, F_GROUPED, F_LEARNED = 1, 2, 4
MAX_TUPLES = 1 << 20


def noalloc_fn(sjbmod):
    f = sjbmod.lib().sjb_lazy_join_noalloc
    assert list(f.argtypes) == [ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                          ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    return f


def test_noalloc_matches_async(ctx, sjbmod, torch_mod, orc):
    torch = torch_mod
    f = noalloc_fn(sjbmod)
    stream = torch.cuda.Stream()
    cases = [(MAX_TUPLES, MAX_TUPLES, 1 << 22, 11), (70001, 50003, 1 << 12, 12)]
    for nL, nR, key_range, seed in cases:
        print(f"case noalloc nL={nL} nR={nR} range={key_range}", flush=True)
        R = sjbmod.generate(nL, sjbmod.DIST_UNIFORM, key_range, 0.0, seed=seed)
        S = sjbmod.generate(nR, sjbmod.DIST_UNIFORM, key_range, 0.0, seed=seed + 100)
        dR, dS = to_dev(torch, R), to_dev(torch, S)
        torch.cuda.synchronize()
        for alg in (HJ, SMJ):
            for flags in (F_C2, F_GROUPED):
                got = torch.zeros(32, dtype=torch.uint8, device="cuda")
                ref = torch.zeros(32, dtype=torch.uint8, device="cuda")
                torch.cuda.synchronize()
                rc = f(alg, flags, dR.data_ptr(), nL, dS.data_ptr(), nR, got.data_ptr(), stream.cuda_stream)
                assert rc == 0, (alg, flags, rc)
                stream.synchronize()
                ctx.lazy_join_async(alg, flags, dR, nL, dS, nR, ref.data_ptr())
                ctx.sync()
                g = got.cpu().numpy().view(sjbmod.WRES_DTYPE)[0]
                r = ref.cpu().numpy().view(sjbmod.WRES_DTYPE)[0]
                assert (int(g["count"]), int(g["checksum1"])) == (int(r["count"]), int(r["checksum1"])), (alg, flags)
                if flags == F_C2:
                    assert int(g["checksum2"]) == int(r["checksum2"]), alg
        if nL < MAX_TUPLES:
            o, _ = orc.lazy_hj(R, S, 1 << 12)
            assert (int(g["count"]), int(g["checksum1"])) == (o.count, o.c1)


def test_noalloc_empty_side(sjbmod, torch_mod):
    torch = torch_mod
    f = noalloc_fn(sjbmod)
    R = sjbmod.generate(1000, sjbmod.DIST_UNIFORM, 100, 0.0, seed=5)
    dR = to_dev(torch, R)
    out = torch.full((32,), 0xAB, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    s = torch.cuda.current_stream().cuda_stream
    assert f(HJ, F_C2, dR.data_ptr(), 1000, None, 0, out.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert not out.cpu().numpy().any()


def test_noalloc_rejects(sjbmod, torch_mod):
    torch = torch_mod
    f = noalloc_fn(sjbmod)
    d = torch.zeros(64, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    p = d.data_ptr()
    assert f(HJ, 0, p, MAX_TUPLES + 1, p, 10, p, s) == -2   # SJB_E_CAPACITY
    assert f(SMJ, 0, p, 10, p, MAX_TUPLES + 1, p, s) == -2
    assert f(NLJ, 0, p, 1, p, 1, p, s) == -5                # SJB_E_UNSUPPORTED
    assert f(HJ, F_LEARNED, p, 1, p, 1, p, s) == -5
    assert f(HJ, 0, p, 1, p, 1, None, s) == -1              # SJB_E_INVALID
