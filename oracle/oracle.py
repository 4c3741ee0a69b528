"""ctypes binding of the CPU oracle (oracle/libsjb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py — as the checker or the timed CPU baseline,
never as the product path.  See sjb_oracle.h for what each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsjb_oracle.so")

TUPLE_DTYPE = np.dtype([("key", "<u4"), ("val", "<u4"), ("ts", "<u8")])
PAIR_DTYPE = np.dtype([("lkey", "<u4"), ("rkey", "<u4"), ("lval", "<u4"), ("rval", "<u4"),
                       ("lts", "<u8"), ("rts", "<u8")])


class Result(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("c1", ctypes.c_uint64), ("c2", ctypes.c_uint64),
                ("pairs", ctypes.c_uint64)]

    def triple(self):
        return (self.count, self.c1, self.c2)


_lib = None
_p, _u64, _i, _u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
_R = ctypes.POINTER(Result)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_mix64": (_u64, [_u64]),
            "or_hash_a": (_u64, [_u32, _u32]),
            "or_hash_b": (_u64, [_u32, _u32]),
            "or_pair_hash": (_u64, [_u32, _u32, _u32]),
            "or_lazy_hj": (_i, [_p, _u64, _p, _u64, _u64, _i, _R, _p, _u64]),
            "or_lazy_smj": (_i, [_p, _u64, _p, _u64, _i, _R, _p, _u64]),
            "or_lazy_nlj": (_i, [_p, _u64, _p, _u64, _R, _p, _u64]),
            "or_lazy_batch": (_i, [_i, _p, _p, _p, _p, _u32, _u64, _i, _p]),
            "or_eager_new": (_p, [_u64, _u64]),
            "or_eager_free": (None, [_p]),
            "or_eager_clear": (None, [_p]),
            "or_eager_execute": (_i, [_p, _p, _u64, _i, _i, _R, _p, _u64]),
            "or_sliding": (_i, [_p, _p, _u64, _u64, _u64, _p]),
            "or_bench_lazy_hj": (ctypes.c_double, [_p, _p, _u64, _u64, _u64, _i, _R]),
            "or_bench_lazy_hj_inter": (ctypes.c_double, [_p, _p, _u64, _u64, _u64, _i, _R]),
            "or_bench_slide": (ctypes.c_double, [_p, _p, _p, _p, _u64, _u64, _p]),
            "or_mlp_grad": (ctypes.c_double, [_p, _p, _p, _u32, _p]),
            "or_sliding_var": (_i, [_p, _p, _p, _p, _u64, _u64, _p]),
            "or_mlp_loss": (ctypes.c_double, [_p, _p, _p, _u32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


def _pairs_buf(want, cap):
    return np.zeros(cap, dtype=PAIR_DTYPE) if want else None


def lazy_hj(R, S, D, want_c2=True, want_pairs=False, cap=None):
    cap = cap if cap is not None else (len(R) * len(S) if want_pairs else 0)
    pairs = _pairs_buf(want_pairs, cap)
    r = Result()
    rc = lib().or_lazy_hj(_ptr(R), len(R), _ptr(S), len(S), D, int(want_c2), ctypes.byref(r),
                          _ptr(pairs), cap if want_pairs else 0)
    assert rc == 0
    return r, (pairs[: r.pairs] if want_pairs else None)


def lazy_smj(R, S, want_c2=False, want_pairs=False, cap=None):
    cap = cap if cap is not None else (len(R) * len(S) if want_pairs else 0)
    pairs = _pairs_buf(want_pairs, cap)
    r = Result()
    rc = lib().or_lazy_smj(_ptr(R), len(R), _ptr(S), len(S), int(want_c2), ctypes.byref(r),
                           _ptr(pairs), cap if want_pairs else 0)
    assert rc == 0
    return r, (pairs[: r.pairs] if want_pairs else None)


def lazy_nlj(R, S, want_pairs=False, cap=None):
    cap = cap if cap is not None else (len(R) * len(S) if want_pairs else 0)
    pairs = _pairs_buf(want_pairs, cap)
    r = Result()
    rc = lib().or_lazy_nlj(_ptr(R), len(R), _ptr(S), len(S), ctypes.byref(r), _ptr(pairs),
                           cap if want_pairs else 0)
    assert rc == 0
    return r, (pairs[: r.pairs] if want_pairs else None)


class Eager:
    def __init__(self, D, max_window):
        self.h = lib().or_eager_new(D, max_window)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_eager_free(self.h)
            self.h = None

    def clear(self):
        lib().or_eager_clear(self.h)

    def execute(self, batch, is_left, want_c2=True):
        r = Result()
        rc = lib().or_eager_execute(self.h, _ptr(batch), len(batch), int(is_left), int(want_c2),
                                    ctypes.byref(r), None, 0)
        assert rc == 0, rc
        return r


def sliding(R, S, W, slide):
    n = len(R)
    nw = n // slide
    res = (Result * max(nw, 1))()
    rc = lib().or_sliding(_ptr(R), _ptr(S), n, W, slide, ctypes.cast(res, ctypes.c_void_p))
    assert rc == 0
    return [res[i].triple() for i in range(nw)]


def sliding_var(R, endR, S, endS, W):
    """(count, C1) of the last-W-tuples window of each side after every event."""
    endR = np.ascontiguousarray(endR, dtype=np.uint64)
    endS = np.ascontiguousarray(endS, dtype=np.uint64)
    n = len(endR)
    res = (Result * max(n, 1))()
    rc = lib().or_sliding_var(_ptr(R), endR.ctypes.data, _ptr(S), endS.ctypes.data, n, W,
                              ctypes.cast(res, ctypes.c_void_p))
    assert rc == 0
    return [(res[i].count, res[i].c1) for i in range(n)]


def bench_slide(R, endR, S, endS, W):
    """Seconds and per-event (count, C1) of the incremental CPU sliding join."""
    endR = np.ascontiguousarray(endR, dtype=np.uint64)
    endS = np.ascontiguousarray(endS, dtype=np.uint64)
    n = len(endR)
    res = (Result * max(n, 1))()
    secs = lib().or_bench_slide(_ptr(R), endR.ctypes.data, _ptr(S), endS.ctypes.data, n, W,
                                ctypes.cast(res, ctypes.c_void_p))
    return secs, [(res[i].count, res[i].c1) for i in range(n)]


def bench_lazy_hj(R, S, W, D, threads, inter=False):
    """Seconds and totals of the timed CPU hash join over the tumbling
    windows of R and S: INTRA (1 joiner x `threads` probe threads) or INTER
    (`threads` joiners taking whole windows)."""
    tot = Result()
    fn = lib().or_bench_lazy_hj_inter if inter else lib().or_bench_lazy_hj
    secs = fn(_ptr(R), _ptr(S), min(len(R), len(S)), W, D, threads, ctypes.byref(tot))
    return secs, tot


def sort_pairs(p):
    """Canonical order of a pair list (by every field)."""
    return np.sort(p, order=["lkey", "lval", "rval", "lts", "rts", "rkey"])


MLP_PARAMS = 1153


def mlp_grad(P, x, t):
    """float64 loss and gradient of the learned partitioner's MLP root."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.zeros(MLP_PARAMS, dtype=np.float64)
    loss = lib().or_mlp_grad(P.ctypes.data, x.ctypes.data, t.ctypes.data, len(x), g.ctypes.data)
    return loss, g


def mlp_loss(P, x, t):
    P = np.ascontiguousarray(P, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    return lib().or_mlp_loss(P.ctypes.data, x.ctypes.data, t.ctypes.data, len(x))
