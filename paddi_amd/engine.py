"""Python front-end of the discord engine: a thin ctypes layer over the C ABI of
``libpaddi_hip.so`` (declared in include/paddi.h).

It mirrors the reference's runner interface -- series, minL..maxL, top-k in; a discord
set of (index, length, squared NN distance) out (src/main.cu:28-33, 236-240,
src/PALMAD_types.hpp:7-11) -- and its per-segment contracts:

  Context.compute_statistics   compute_statistics / compute_statistics_cpu
                               (src/preprocessing.cuh:68, src/cpu_kernels/cpu_kernels.cpp:7)
  Context.range_discords       find_discords_cpu (src/cpu_kernels/cpu_kernels.cpp:33)
  Context.discords             the PALMAD length loop (src/main.cu:136-233)

There is no CPU fallback: if the HIP library is missing or no device is present, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpaddi_hip.so")

PLAIN, BATCHED, LEARNED = 0, 1, 2
MODES = {"plain": PLAIN, "batched": BATCHED, "learned": LEARNED}


class PaddiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Discord(C.Structure):
    _fields_ = [("index", C.c_int64), ("length", C.c_int32), ("series", C.c_int32),
                ("nn_dist2", C.c_float), ("range2", C.c_float)]


class Opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("max_attempts", C.c_int32), ("flags", C.c_int32),
                ("reserved", C.c_int32), ("model_path", C.c_char_p)]


class Stats(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("survivors", C.c_int64),
                ("exact_refines", C.c_int64), ("pairs", C.c_int64), ("seconds", C.c_double),
                ("kernel_ms", C.c_double * 4), ("launches", C.c_int64 * 4),
                ("live_pairs", C.c_int64 * 2), ("scan_pairs", C.c_int64 * 2)]

    KERNELS = ("derive", "select", "refine", "exact")

    def as_dict(self):
        return dict(attempts=self.attempts, survivors=self.survivors,
                    exact_refines=self.exact_refines, pairs=self.pairs, seconds=self.seconds,
                    kernel_ms=dict(zip(self.KERNELS, list(self.kernel_ms))),
                    launches=dict(zip(self.KERNELS, list(self.launches))),
                    scan_pairs=dict(select=self.scan_pairs[0], refine=self.scan_pairs[1]),
                    live_pairs=dict(select=self.live_pairs[0], refine=self.live_pairs[1]))


class Tunables(C.Structure):
    _fields_ = [("select_window", C.c_int64), ("refine_chunk", C.c_int64),
                ("min_chunk", C.c_int64), ("refine_items", C.c_int64),
                ("witness_tiles", C.c_int32), ("speculate", C.c_int32),
                ("use_graph", C.c_int32), ("prev_shift", C.c_int32),
                ("bootstrap", C.c_int32), ("bound_candidates", C.c_int32),
                ("direct_budget", C.c_int32), ("direct_rows", C.c_int32),
                ("device_lengths", C.c_int32), ("device_rounds", C.c_int32),
                ("refine_rounds", C.c_int32), ("seed_witnesses", C.c_int32),
                ("refine_order", C.c_int32)]
    NAMES = ("select_window", "refine_chunk", "min_chunk", "refine_items", "witness_tiles",
             "speculate", "use_graph", "prev_shift", "bootstrap", "bound_candidates",
             "direct_budget", "direct_rows", "device_lengths", "device_rounds",
             "refine_rounds", "seed_witnesses", "refine_order")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.NAMES}


TUNED_PATH = os.path.join(_HERE, "model", "tuned.json")


def tuned_key(minL, arch="gfx950", nseries=1):
    """Autotuner table key: device, length bucket (m < 128, < 256, < 512, >= 512) and, for
    calls of several series, "/batch" (one long series and many short ones scan differently)."""
    b = 64 if minL < 128 else 128 if minL < 256 else 256 if minL < 512 else 512
    return f"{arch}/m{b}" + ("/batch" if nseries > 1 else "")


EXPORTS = ("paddi_get_tunables", "paddi_set_tunables", "paddi_status_str", "paddi_version", "paddi_create", "paddi_destroy",
           "paddi_compute_statistics", "paddi_range_discords", "paddi_discords",
           "paddi_discords_batch", "paddi_series_upload", "paddi_series_free",
           "paddi_discords_series", "paddi_has_mode", "paddi_mlp_sizes", "paddi_features",
           "paddi_mlp_forward", "paddi_mlp_gradient", "paddi_mlp_train",
           "paddi_range_discords_noalloc", "paddi_set_exclusion", "paddi_get_exclusion")

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libpaddi_hip.so and declare the C ABI.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    P, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.paddi_status_str.argtypes = [i32]
    L.paddi_status_str.restype = C.c_char_p
    L.paddi_version.restype = C.c_char_p
    L.paddi_create.argtypes = [i32, C.POINTER(P)]
    L.paddi_destroy.argtypes = [P]
    L.paddi_destroy.restype = None
    L.paddi_compute_statistics.argtypes = [P, P, i64, i32, P, P]
    L.paddi_range_discords.argtypes = [P, P, i64, i32, C.c_float, P, P]
    L.paddi_range_discords_noalloc.argtypes = [P, P, i64, i32, C.c_float, P, P]
    L.paddi_discords.argtypes = [P, P, i64, i32, i32, i32, i32, C.POINTER(Opts), P, i64,
                                 C.POINTER(i64), C.POINTER(Stats)]
    L.paddi_discords_batch.argtypes = [P, P, P, i32, i32, i32, i32, i32, C.POINTER(Opts), P,
                                       i64, C.POINTER(i64), C.POINTER(Stats)]
    L.paddi_series_upload.argtypes = [P, P, i64, C.POINTER(P)]
    L.paddi_series_free.argtypes = [P]
    L.paddi_series_free.restype = None
    L.paddi_discords_series.argtypes = [P, P, i32, i32, i32, i32, i32, C.POINTER(Opts), P, i64,
                                        C.POINTER(i64), C.POINTER(Stats)]
    L.paddi_has_mode.argtypes = [i32]
    L.paddi_has_mode.restype = i32
    L.paddi_get_tunables.argtypes = [P, C.POINTER(Tunables)]
    L.paddi_set_tunables.argtypes = [P, C.POINTER(Tunables)]
    L.paddi_set_exclusion.argtypes = [P, i32]
    L.paddi_get_exclusion.argtypes = [P, C.POINTER(i32)]
    L.paddi_mlp_sizes.argtypes = [P, P, P]
    L.paddi_features.argtypes = [P, P, i32, i32, i32, P]
    L.paddi_mlp_forward.argtypes = [P, P, P, P, P, i32, P]
    L.paddi_mlp_gradient.argtypes = [P, P, P, P, P, P, i32, P, P]
    L.paddi_mlp_train.argtypes = [P, P, P, P, P, P, i32, i32, C.c_float, C.c_float, P]
    _lib = L
    return L


def mlp_sizes():
    L = load_library()
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    L.paddi_mlp_sizes(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def _f32(T):
    return np.ascontiguousarray(T, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class DiscordRecord:
    index: int
    length: int
    nn_dist2: float
    range2: float
    series: int = 0


class SeriesHandle:
    """A series uploaded to HBM (paddi_series_upload)."""

    def __init__(self, ctx: "Context", T):
        self._ctx = ctx
        self.n = len(T)
        self._h = C.c_void_p()
        T = _f32(T)
        ctx._check(ctx._L.paddi_series_upload(ctx._c, _ptr(T), len(T), C.byref(self._h)),
                   "paddi_series_upload")

    def free(self):
        if self._h:
            self._ctx._L.paddi_series_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        self._L = load_library()
        self._c = C.c_void_p()
        self._check(self._L.paddi_create(device, C.byref(self._c)), "paddi_create")

    def _check(self, st, what):
        if st != 0:
            raise PaddiError(st, f"{what}: {self._L.paddi_status_str(st).decode()}")

    def close(self):
        if self._c:
            self._L.paddi_destroy(self._c)
            self._c = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- schedule tunables (autotuner) ---------------------------------------------

    def get_tunables(self):
        t = Tunables()
        self._check(self._L.paddi_get_tunables(self._c, C.byref(t)), "paddi_get_tunables")
        return t.as_dict()

    def set_tunables(self, **kw):
        t = Tunables()
        self._check(self._L.paddi_get_tunables(self._c, C.byref(t)), "paddi_get_tunables")
        for k, v in kw.items():
            if k not in Tunables.NAMES:
                raise KeyError(k)
            setattr(t, k, int(v))
        self._check(self._L.paddi_set_tunables(self._c, C.byref(t)), "paddi_set_tunables")

    # -- exclusion rule ----------------------------------------------------------------

    EXCLUSION = {"cpu": 0, "symmetric": 1}

    def set_exclusion(self, rule):
        """"cpu": pair iff |i-1-j| >= m (src/cpu_kernels/cpu_kernels.cpp:57, the default);
        "symmetric": iff |i-j| >= m-1 (src/PD3_kernels.cuh:70)."""
        code = self.EXCLUSION[rule] if isinstance(rule, str) else int(rule)
        self._check(self._L.paddi_set_exclusion(self._c, code), "paddi_set_exclusion")

    def get_exclusion(self):
        v = C.c_int32()
        self._check(self._L.paddi_get_exclusion(self._c, C.byref(v)), "paddi_get_exclusion")
        return {c: n for n, c in self.EXCLUSION.items()}[v.value]

    def apply_tuned(self, minL, path=TUNED_PATH, nseries=1):
        """Apply the autotuner's entry for this length range and call shape (one series or
        several; the other shape's entry if only that one exists), if the table has one."""
        import json
        if not os.path.exists(path):
            return None
        with open(path) as f:
            table = json.load(f)
        ent = table.get(tuned_key(minL, nseries=nseries)) or \
            table.get(tuned_key(minL, nseries=1 if nseries > 1 else 2))
        if ent:
            self.set_tunables(**ent["tunables"])
        return ent

    # -- per-segment contracts ---------------------------------------------------

    def compute_statistics(self, T, m):
        T = _f32(T)
        N = len(T) - m + 1
        mu, sd = np.empty(max(N, 0), np.float32), np.empty(max(N, 0), np.float32)
        self._check(self._L.paddi_compute_statistics(self._c, _ptr(T), len(T), m, _ptr(mu),
                                                     _ptr(sd)), "paddi_compute_statistics")
        return mu, sd

    def range_discords(self, T, m, r2):
        T = _f32(T)
        N = len(T) - m + 1
        flag = np.empty(max(N, 0), np.int32)
        nn = np.empty(max(N, 0), np.float32)
        self._check(self._L.paddi_range_discords(self._c, _ptr(T), len(T), m, C.c_float(r2),
                                                 _ptr(flag), _ptr(nn)), "paddi_range_discords")
        return flag, nn

    def range_discords_noalloc(self, T, m, r2):
        """range_discords through the entry point that allocates nothing (its buffers are a
        fixed device array; series of up to 300,000 points)."""
        T = _f32(T)
        N = len(T) - m + 1
        flag = np.empty(max(N, 0), np.int32)
        nn = np.empty(max(N, 0), np.float32)
        self._check(self._L.paddi_range_discords_noalloc(self._c, _ptr(T), len(T), m,
                                                         C.c_float(r2), _ptr(flag), _ptr(nn)),
                    "paddi_range_discords_noalloc")
        return flag, nn

    # -- runner API --------------------------------------------------------------

    @staticmethod
    def _opts(mode, max_attempts, model_path):
        return Opts(MODES[mode] if isinstance(mode, str) else int(mode), max_attempts, 0, 0,
                    model_path.encode() if model_path else None)

    @staticmethod
    def _records(buf, count):
        return [DiscordRecord(int(d.index), int(d.length), float(d.nn_dist2), float(d.range2),
                              int(d.series)) for d in buf[:count]]

    def discords(self, T, minL, maxL, step=1, top_k=1, mode="plain", max_attempts=64,
                 model_path=None, with_stats=False):
        return self.discords_batch([T], minL, maxL, step, top_k, mode, max_attempts, model_path,
                                   with_stats)

    def discords_batch(self, series, minL, maxL, step=1, top_k=1, mode="plain",
                       max_attempts=64, model_path=None, with_stats=False):
        arrs = [_f32(T) for T in series]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        ns = np.array([len(a) for a in arrs], np.int64)
        nlen = (maxL - minL) // step + 1
        cap = nlen * top_k * len(arrs)
        buf = (Discord * max(cap, 1))()
        cnt = C.c_int64(0)
        st = Stats()
        o = self._opts(mode, max_attempts, model_path)
        self._check(self._L.paddi_discords_batch(self._c, ptrs, _ptr(ns), len(arrs), minL, maxL,
                                                 step, top_k, C.byref(o), buf, cap,
                                                 C.byref(cnt), C.byref(st)),
                    "paddi_discords_batch")
        recs = self._records(buf, cnt.value)
        return (recs, st.as_dict()) if with_stats else recs

    def upload(self, T) -> SeriesHandle:
        return SeriesHandle(self, T)

    # -- learned range proposal ---------------------------------------------------

    def features(self, handles, m, top_k=1):
        nf, _, _ = mlp_sizes()
        hs = (C.c_void_p * len(handles))(*[h._h.value for h in handles])
        out = np.empty((len(handles), nf), np.float32)
        self._check(self._L.paddi_features(self._c, hs, len(handles), m, top_k, _ptr(out)),
                    "paddi_features")
        return out

    def mlp_forward(self, params, fmu, fsd, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty(len(X), np.float32)
        args = [np.ascontiguousarray(a, np.float32) for a in (params, fmu, fsd)]
        self._check(self._L.paddi_mlp_forward(self._c, *[_ptr(a) for a in args], _ptr(X),
                                              len(X), _ptr(Y)), "paddi_mlp_forward")
        return Y

    def mlp_gradient(self, params, fmu, fsd, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        args = [np.ascontiguousarray(a, np.float32) for a in (params, fmu, fsd)]
        g = np.empty(len(args[0]), np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self._L.paddi_mlp_gradient(self._c, *[_ptr(a) for a in args], _ptr(X),
                                               _ptr(T), len(X), _ptr(g), _ptr(loss)),
                    "paddi_mlp_gradient")
        return g, float(loss[0])

    def mlp_train(self, params, fmu, fsd, X, T, steps, lr=1e-3, weight_decay=0.0):
        p = np.array(params, np.float32, copy=True)
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        fmu = np.ascontiguousarray(fmu, np.float32)
        fsd = np.ascontiguousarray(fsd, np.float32)
        losses = np.empty(steps, np.float32)
        self._check(self._L.paddi_mlp_train(self._c, _ptr(p), _ptr(fmu), _ptr(fsd), _ptr(X),
                                            _ptr(T), len(X), steps, C.c_float(lr),
                                            C.c_float(weight_decay), _ptr(losses)),
                    "paddi_mlp_train")
        return p, losses

    def discords_series(self, handles, minL, maxL, step=1, top_k=1, mode="plain",
                        max_attempts=64, model_path=None, with_stats=False):
        hs = (C.c_void_p * len(handles))(*[h._h.value for h in handles])
        nlen = (maxL - minL) // step + 1
        cap = nlen * top_k * len(handles)
        buf = (Discord * max(cap, 1))()
        cnt = C.c_int64(0)
        st = Stats()
        o = self._opts(mode, max_attempts, model_path)
        self._check(self._L.paddi_discords_series(self._c, hs, len(handles), minL, maxL, step,
                                                  top_k, C.byref(o), buf, cap, C.byref(cnt),
                                                  C.byref(st)), "paddi_discords_series")
        recs = self._records(buf, cnt.value)
        return (recs, st.as_dict()) if with_stats else recs
