"""ctypes binding of libsjb_hip.so (the C ABI declared in include/sjb.h).

Python callers (tests, bench.py, __graft_entry__) reach the HIP operators only
through this module.  It never falls back to anything: if the shared library
is missing or fails to load, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.environ.get("SJB_LIB") or os.path.join(PKG, "lib", "libsjb_hip.so")

# enums (include/sjb/windowing/BaseWindowing.cuh:9-40 values)
ALG = {"EMPTY": 0, "NLJ": 1, "HJ": 2, "SMJ": 3}
OUT = {"NoOutput": 0, "CountKernel": 1, "EstimatedSelectivity": 2, "Bitmap": 3, "Atomic": 4}
F_C2, F_GROUPED, F_LEARNED, F_TIMING, F_OVERLAP, F_NO_REFIT, F_PEEL = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20, 0x40
DIST_UNIFORM, DIST_ZIPF = 0, 1

E = {
    0: "SJB_OK", -1: "SJB_E_INVALID", -2: "SJB_E_CAPACITY", -3: "SJB_E_NOMEM", -4: "SJB_E_HIP",
    -5: "SJB_E_UNSUPPORTED", -6: "SJB_E_NODEVICE", -7: "SJB_E_BUDGET", -8: "SJB_E_TIMEOUT",
    -9: "SJB_E_STATE", -10: "SJB_E_MISMATCH",
}

TUPLE_DTYPE = np.dtype([("key", "<u4"), ("val", "<u4"), ("ts", "<u8")])
RESULT_DTYPE = np.dtype([("lkey", "<u4"), ("rkey", "<u4"), ("lval", "<u4"), ("rval", "<u4"),
                         ("lts", "<u8"), ("rts", "<u8")])
WRES_DTYPE = np.dtype([("count", "<u8"), ("checksum1", "<u8"), ("checksum2", "<u8"),
                       ("pairs", "<u8")])


class SjbError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {E.get(code, code)}")


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "result_count", "build_time_us", "probe_time_us", "sort_time_us", "merge_time_us",
        "checksum1", "checksum2", "pairs_written", "device_ns")] + [
        ("c2_valid", ctypes.c_uint32), ("fanout_log2", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Options(ctypes.Structure):
    _fields_ = [("max_window_tuples", ctypes.c_uint64), ("num_distinct_keys", ctypes.c_uint64),
                ("batch_size", ctypes.c_uint64), ("hbm_budget_bytes", ctypes.c_uint64),
                ("fanout_log2", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class BatchTune(ctypes.Structure):
    _fields_ = [("fanout_bias", ctypes.c_int32), ("chunk_log2", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("grab", ctypes.c_uint32),
                ("expand", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("ms", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_windows", ctypes.c_uint32), ("n_chunks", ctypes.c_uint32),
                ("n_items", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("tune", BatchTune),
                ("part_kind", ctypes.c_uint32), ("drift_fallback", ctypes.c_uint32),
                ("drift_refit", ctypes.c_uint32), ("n_heavy", ctypes.c_uint32)]


class RmiConfig(ctypes.Structure):
    _fields_ = [("root", ctypes.c_uint32), ("steps", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("heavy_min", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class RmiReport(ctypes.Structure):
    _fields_ = [("loss_init", ctypes.c_double), ("loss_final", ctypes.c_double),
                ("fit_ms", ctypes.c_double), ("imbalance_learned", ctypes.c_double),
                ("imbalance_radix", ctypes.c_double), ("n_heavy", ctypes.c_uint32),
                ("leaves_used", ctypes.c_uint32), ("n_sample", ctypes.c_uint32),
                ("fallback", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class PartitionReport(ctypes.Structure):
    _fields_ = [("fanout_log2", ctypes.c_uint32), ("kind", ctypes.c_uint32),
                ("max_build", ctypes.c_uint64), ("max_probe", ctypes.c_uint64),
                ("mean_build", ctypes.c_double), ("imbalance", ctypes.c_double),
                ("fixup_frac", ctypes.c_double), ("heavy_left", ctypes.c_uint64),
                ("heavy_right", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SlideInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("window", "max_batch", "table_slots", "device_bytes",
                                               "events", "rebuilds", "tuples_left", "tuples_right")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


RMI_LINEAR, RMI_MLP = 0, 1
MLP_PARAMS = 1153

# exported symbols and their signatures: (restype, argtypes)
_u64, _u32, _i, _p, _d = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_double
SIGNATURES = {
    "sjb_abi_version": (_i, []),
    "sjb_strerror": (ctypes.c_char_p, [_i]),
    "sjb_open": (_i, [_i, ctypes.POINTER(Options), ctypes.POINTER(_p)]),
    "sjb_close": (_i, [_p]),
    "sjb_stream": (_p, [_p]),
    "sjb_sync": (_i, [_p]),
    "sjb_device_bytes": (_u64, [_p]),
    "sjb_host_alloc": (_i, [_u64, ctypes.POINTER(_p)]),
    "sjb_host_free": (_i, [_p]),
    "sjb_lazy_join": (_i, [_p, _i, _i, _u32, _p, _u64, _p, _u64, _p, _u64, ctypes.POINTER(Stats)]),
    "sjb_lazy_execute": (_i, [_p, _i, _i, _u32, _p, _p, _u64, _p, _p, _u64, _u64, _p, _u64,
                              ctypes.POINTER(Stats)]),
    "sjb_lazy_join_async": (_i, [_p, _i, _u32, _p, _u64, _p, _u64, _p]),
    "sjb_lazy_join_noalloc": (_i, [_i, _u32, _p, _u64, _p, _u64, _p, _p]),
    "sjb_lazy_join_batch": (_i, [_p, _i, _u32, _p, _p, _p, _p, _u32, _p]),
    "sjb_batch_plan_create": (_i, [_p, _i, _u32, _p, _p, _p, _p, _u32, _p, ctypes.POINTER(_p)]),
    "sjb_batch_plan_launch": (_i, [_p]),
    "sjb_batch_plan_info": (_i, [_p, ctypes.POINTER(BatchInfo)]),
    "sjb_batch_plan_destroy": (_i, [_p]),
    "sjb_batch_autotune": (_i, [_p, _i, _u32, _p, _p, _p, _p, _u32, _u32, ctypes.POINTER(BatchTune),
                                ctypes.POINTER(_u32)]),
    "sjb_batch_get_tune": (_i, [_p, ctypes.POINTER(BatchTune)]),
    "sjb_batch_set_tune": (_i, [_p, ctypes.POINTER(BatchTune)]),
    "sjb_timing_ring": (_i, [_p, _u32]),
    "sjb_timing_read": (_i, [_p, _p, _p, _u32, ctypes.POINTER(_u32)]),
    "sjb_rmi_fit": (_i, [_p, _p, _u64, ctypes.POINTER(RmiConfig), ctypes.POINTER(RmiReport)]),
    "sjb_rmi_clear": (_i, [_p]),
    "sjb_mlp_grad": (_i, [_p, _p, _p, _p, _u32, _p, ctypes.POINTER(_d)]),
    "sjb_partition_report": (_i, [_p, _i, _u32, _p, _u64, _p, _u64, ctypes.POINTER(PartitionReport)]),
    "sjb_generate_mix": (_i, [_p, _u64, _u64, _d, _d, _u64, _u64]),
    "sjb_slide_create": (_i, [_p, _u64, _u64, ctypes.POINTER(_p)]),
    "sjb_slide_reset": (_i, [_p]),
    "sjb_slide_push": (_i, [_p, _p, _u64, _p, _u64, _p]),
    "sjb_slide_serve": (_i, [_p, _p, _p, _p, _p, _u32, _p, _p, _p, ctypes.POINTER(_d)]),
    "sjb_slide_info": (_i, [_p, ctypes.POINTER(SlideInfo)]),
    "sjb_slide_destroy": (_i, [_p]),
    "sjb_eager_execute": (_i, [_p, _i, _i, _u32, _p, _u64, _i, _p, _u64, ctypes.POINTER(Stats)]),
    "sjb_eager_clear": (_i, [_p]),
    "sjb_eager_end_of_stream": (_i, [_p]),
    "sjb_pk_execute": (_i, [_p, _p, _u64, _i, ctypes.POINTER(Stats)]),
    "sjb_pk_clear": (_i, [_p]),
    "sjb_pk_stop": (_i, [_p]),
    "sjb_pk_launches": (_u64, [_p]),
    "sjb_generate": (_i, [_p, _u64, _i, _u64, _d, _u64, _u64]),
    "sjb_mix64": (_u64, [_u64]),
    "sjb_hash_a": (_u64, [_u32, _u32]),
    "sjb_hash_b": (_u64, [_u32, _u32]),
    "sjb_pair_hash": (_u64, [_u32, _u32, _u32]),
}

_lib = None


def lib():
    """Loads libsjb_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsjb_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise SjbError(rc, what)
    return rc


def ptr(a):
    """Raw address of a numpy array or a torch tensor."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def generate_mix(n, window, zipf_frac=0.3, theta=1.2, seed=123, ts_base=0):
    """Drifting 8-component Gaussian mixture over u32 keys mixed with Zipf keys."""
    out = np.empty(n, dtype=TUPLE_DTYPE)
    check(lib().sjb_generate_mix(ptr(out), n, window, zipf_frac, theta, seed, ts_base),
          "sjb_generate_mix")
    return out


def generate(n, dist=DIST_UNIFORM, key_range=1 << 16, theta=0.0, seed=123, ts_base=0):
    out = np.empty(n, dtype=TUPLE_DTYPE)
    check(lib().sjb_generate(ptr(out), n, dist, key_range, theta, seed, ts_base), "sjb_generate")
    return out


class Context:
    """One sjb_ctx: a HIP stream plus the device buffers of one window size."""

    def __init__(self, device=0, max_window=1 << 22, distinct_keys=1000, batch_size=8192,
                 hbm_budget=0, fanout_log2=0, flags=0):
        o = Options(max_window, distinct_keys, batch_size, hbm_budget, fanout_log2, flags)
        h = ctypes.c_void_p()
        self._plans = []
        check(lib().sjb_open(device, ctypes.byref(o), ctypes.byref(h)), "sjb_open")
        self.h = h

    def close(self):
        if self.h:
            for plan in self._plans:
                plan.destroy()
            self._plans = []
            lib().sjb_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().sjb_stream(self.h)

    def sync(self):
        check(lib().sjb_sync(self.h), "sjb_sync")

    def device_bytes(self):
        return lib().sjb_device_bytes(self.h)

    def lazy_join(self, alg, method, flags, left, n_left, right, n_right, out=None, out_cap=0,
                  raise_on_error=True):
        st = Stats()
        rc = lib().sjb_lazy_join(self.h, alg, method, flags, ptr(left), n_left, ptr(right), n_right,
                                 ptr(out), out_cap, ctypes.byref(st))
        if raise_on_error:
            check(rc, "sjb_lazy_join")
        return (rc, st) if not raise_on_error else st

    def lazy_execute(self, alg, method, flags, left_ring, left_idx, right_ring, right_idx,
                     batch_size, out=None, out_cap=0):
        st = Stats()
        li = np.ascontiguousarray(left_idx, dtype=np.uint64)
        ri = np.ascontiguousarray(right_idx, dtype=np.uint64)
        rc = lib().sjb_lazy_execute(self.h, alg, method, flags, ptr(left_ring), ptr(li), len(li),
                                    ptr(right_ring), ptr(ri), len(ri), batch_size, ptr(out),
                                    out_cap, ctypes.byref(st))
        return rc, st

    def lazy_join_async(self, alg, flags, left, n_left, right, n_right, d_result):
        check(lib().sjb_lazy_join_async(self.h, alg, flags, ptr(left), n_left, ptr(right), n_right,
                                        d_result), "sjb_lazy_join_async")

    def lazy_join_batch(self, alg, flags, left, left_off, right, right_off):
        lo = np.ascontiguousarray(left_off, dtype=np.uint64)
        ro = np.ascontiguousarray(right_off, dtype=np.uint64)
        nw = len(lo) - 1
        res = np.zeros(nw, dtype=WRES_DTYPE)
        check(lib().sjb_lazy_join_batch(self.h, alg, flags, ptr(left), ptr(lo), ptr(right), ptr(ro),
                                        nw, ptr(res)), "sjb_lazy_join_batch")
        return res

    def batch_plan(self, alg, flags, left, left_off, right, right_off, d_results):
        """Creates a captured batch plan; d_results: device array of len(left_off)-1
        sjb_window_result_t (32 B each)."""
        lo = np.ascontiguousarray(left_off, dtype=np.uint64)
        ro = np.ascontiguousarray(right_off, dtype=np.uint64)
        h = ctypes.c_void_p()
        check(lib().sjb_batch_plan_create(self.h, alg, flags, ptr(left), ptr(lo), ptr(right), ptr(ro),
                                          len(lo) - 1, ptr(d_results), ctypes.byref(h)),
              "sjb_batch_plan_create")
        plan = BatchPlan(h, self)
        self._plans.append(plan)
        return plan

    def slide(self, window, max_batch):
        """Creates a sliding-window join state (raises SjbError SJB_E_BUDGET when it
        would not fit the ctx's HBM budget)."""
        h = ctypes.c_void_p()
        check(lib().sjb_slide_create(self.h, window, max_batch, ctypes.byref(h)), "sjb_slide_create")
        sl = Slide(h, self)
        self._plans.append(sl)
        return sl

    def batch_autotune(self, alg, flags, left, left_off, right, right_off, reps=3):
        lo = np.ascontiguousarray(left_off, dtype=np.uint64)
        ro = np.ascontiguousarray(right_off, dtype=np.uint64)
        best, tried = BatchTune(), ctypes.c_uint32()
        check(lib().sjb_batch_autotune(self.h, alg, flags, ptr(left), ptr(lo), ptr(right), ptr(ro),
                                       len(lo) - 1, reps, ctypes.byref(best), ctypes.byref(tried)),
              "sjb_batch_autotune")
        return best, tried.value

    def batch_tune(self, **kw):
        """Returns the ctx's batch launch shape; keyword arguments set fields first."""
        t = BatchTune()
        check(lib().sjb_batch_get_tune(self.h, ctypes.byref(t)), "sjb_batch_get_tune")
        if kw:
            for k, v in kw.items():
                setattr(t, k, v)
            check(lib().sjb_batch_set_tune(self.h, ctypes.byref(t)), "sjb_batch_set_tune")
        return t

    def timing_ring(self, n):
        check(lib().sjb_timing_ring(self.h, n), "sjb_timing_ring")

    def timing_read(self, n):
        """(partition ms, join ms) arrays of the last n timed async windows."""
        a, b = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
        got = ctypes.c_uint32()
        check(lib().sjb_timing_read(self.h, ptr(a), ptr(b), n, ctypes.byref(got)), "sjb_timing_read")
        return a[:got.value], b[:got.value]

    def rmi_fit(self, train, n_train, root=RMI_MLP, steps=400, lr=0.01, seed=3, heavy_min=3):
        cfg = RmiConfig(root, steps, lr, seed, heavy_min, 0)
        rep = RmiReport()
        check(lib().sjb_rmi_fit(self.h, ptr(train), n_train, ctypes.byref(cfg), ctypes.byref(rep)),
              "sjb_rmi_fit")
        return rep

    def rmi_clear(self):
        check(lib().sjb_rmi_clear(self.h), "sjb_rmi_clear")

    def mlp_grad(self, params, x, t):
        P = np.ascontiguousarray(params, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        check(lib().sjb_mlp_grad(self.h, ptr(P), ptr(x), ptr(t), len(x), ptr(g), ctypes.byref(loss)),
              "sjb_mlp_grad")
        return loss.value, g

    def partition_report(self, alg, flags, left, n_left, right, n_right):
        rep = PartitionReport()
        check(lib().sjb_partition_report(self.h, alg, flags, ptr(left), n_left, ptr(right), n_right,
                                         ctypes.byref(rep)), "sjb_partition_report")
        return rep

    def eager_execute(self, alg, method, flags, batch, is_left, out=None, out_cap=0):
        st = Stats()
        rc = lib().sjb_eager_execute(self.h, alg, method, flags, ptr(batch), len(batch),
                                     1 if is_left else 0, ptr(out), out_cap, ctypes.byref(st))
        return rc, st

    def eager_clear(self):
        check(lib().sjb_eager_clear(self.h), "sjb_eager_clear")

    def pk_execute(self, batch, is_left):
        st = Stats()
        rc = lib().sjb_pk_execute(self.h, ptr(batch), len(batch), 1 if is_left else 0, ctypes.byref(st))
        return rc, st

    def pk_clear(self):
        check(lib().sjb_pk_clear(self.h), "sjb_pk_clear")

    def pk_stop(self):
        check(lib().sjb_pk_stop(self.h), "sjb_pk_stop")

    def pk_launches(self):
        return lib().sjb_pk_launches(self.h)


class BatchPlan:
    """A captured batch of windows (sjb_batch_plan_*): launch() replays the graph."""

    def __init__(self, h, ctx):
        self.h = h
        self.ctx = ctx  # plans are destroyed before their ctx closes

    def launch(self):
        if not self.h:
            raise SjbError(-1, "sjb_batch_plan_launch (destroyed plan)")
        check(lib().sjb_batch_plan_launch(self.h), "sjb_batch_plan_launch")

    def info(self):
        i = BatchInfo()
        check(lib().sjb_batch_plan_info(self.h, ctypes.byref(i)), "sjb_batch_plan_info")
        return i

    def destroy(self):
        if self.h:
            lib().sjb_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Slide:
    """Sliding-window join state (sjb_slide_*)."""

    def __init__(self, h, ctx):
        self.h = h
        self.ctx = ctx

    def push(self, left, right):
        r = np.zeros(1, dtype=WRES_DTYPE)
        check(lib().sjb_slide_push(self.h, ptr(left) if len(left) else None, len(left),
                                   ptr(right) if len(right) else None, len(right), ptr(r)),
              "sjb_slide_push")
        return int(r[0]["count"]), int(r[0]["checksum1"])

    def serve(self, left, left_off, right, right_off, arrival_s):
        """Runs the serving loop; returns (results, latency_s, elapsed_s)."""
        lo = np.ascontiguousarray(left_off, dtype=np.uint64)
        ro = np.ascontiguousarray(right_off, dtype=np.uint64)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        n = len(lo) - 1
        lat = np.zeros(n, dtype=np.float64)
        res = np.zeros(n, dtype=WRES_DTYPE)
        el = ctypes.c_double()
        check(lib().sjb_slide_serve(self.h, ptr(left), ptr(lo), ptr(right), ptr(ro), n, ptr(arr), ptr(lat),
                                    ptr(res), ctypes.byref(el)), "sjb_slide_serve")
        return res, lat, el.value

    def reset(self):
        check(lib().sjb_slide_reset(self.h), "sjb_slide_reset")

    def info(self):
        i = SlideInfo()
        check(lib().sjb_slide_info(self.h, ctypes.byref(i)), "sjb_slide_info")
        return i

    def destroy(self):
        if self.h:
            lib().sjb_slide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
