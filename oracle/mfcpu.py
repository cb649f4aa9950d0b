"""ctypes wrapper of the CPU oracle (oracle/libmfcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
as the checker and by bench.py's cpu_baseline leg. The product package
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from molflow_sources_amd._abi import ModelView, TallyLayout, TallyView
from molflow_sources_amd.tallies import OFFSET_KINDS, Tallies

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmfcpu.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, u64, i32, dp = C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(C.c_double)
        L.mfcpu_load.restype = vp
        L.mfcpu_load.argtypes = [C.POINTER(ModelView), C.c_char_p, C.c_size_t]
        L.mfcpu_free.argtypes = [vp]
        L.mfcpu_layout.argtypes = [vp, C.POINTER(TallyLayout)]
        L.mfcpu_facet_offsets.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.c_size_t]
        L.mfcpu_run.argtypes = [vp, u64, u64, u64, u64, C.c_int, C.POINTER(TallyView)]
        L.mfcpu_trace.argtypes = [vp, u64, u64, C.c_uint32, C.c_uint32, C.POINTER(i32), C.POINTER(i32)]
        L.mfcpu_intersect_batch.argtypes = [vp, u64, dp, dp, C.POINTER(i32), i32, C.c_int, C.POINTER(i32), dp, dp, dp]
        L.mfcpu_bvh_stats.argtypes = [vp, u64, dp, dp, C.POINTER(i32), i32, dp, dp]
        L.mfcpu_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.mfcpu_maxwell_cdf.argtypes = [C.c_double, dp, dp, C.c_size_t]
        L.mfcpu_parse_user_moment.argtypes = [C.c_char_p, C.c_double, dp, C.c_size_t, C.POINTER(C.c_size_t)]
        L.mfcpu_integrated_desorption.argtypes = [dp, dp, C.c_size_t, C.c_int, C.c_int, C.c_double, dp, dp,
                                                  C.c_size_t, C.POINTER(C.c_size_t)]
        L.mfcpu_facet_area.restype = C.c_double
        L.mfcpu_facet_area.argtypes = [vp, C.c_uint32]
        L.mfcpu_facet_frame.argtypes = [vp, C.c_uint32, dp]
        L.mfcpu_texture_increments.argtypes = [vp, C.c_uint32, dp, C.c_size_t]
        for n in ("mfcpu_total_desorbed_molecules", "mfcpu_final_outgassing_rate", "mfcpu_latest_moment"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [vp]
        L.mfcpu_scatter_sample.argtypes = [C.c_double, C.c_double, u64, u64, dp, dp]
        L.mfcpu_adjoint_run.argtypes = [vp, u64, u64, u64, C.POINTER(i32), C.c_uint32, dp, C.POINTER(i32),
                                        C.c_uint32, C.c_int, dp, dp, u64]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class CpuModel:
    """A model prepared by the oracle."""

    def __init__(self, model):
        self.model = model
        err = C.create_string_buffer(512)
        self._view = model.view()
        self.h = lib().mfcpu_load(C.byref(self._view), err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().mfcpu_free(self.h)
            self.h = None

    def layout(self) -> TallyLayout:
        L = TallyLayout()
        lib().mfcpu_layout(self.h, C.byref(L))
        return L

    def offsets(self) -> dict:
        out = {}
        n = self.model.n_facets
        for name, k in OFFSET_KINDS.items():
            a = np.zeros(n, dtype=np.int64)
            lib().mfcpu_facet_offsets(self.h, k, a.ctypes.data_as(C.POINTER(C.c_int64)), n)
            out[name] = a
        return out

    def new_tallies(self) -> Tallies:
        return Tallies(self.layout(), self.offsets())

    def run(self, n_des: int, seed: int = 42, first: int = 0, threads: int = 0, max_events: int = 0,
            tallies: Tallies | None = None) -> Tallies:
        t = tallies or self.new_tallies()
        v = t.view()
        rc = lib().mfcpu_run(self.h, seed, first, n_des, max_events, threads, C.byref(v))
        if rc:
            raise RuntimeError(f"mfcpu_run failed ({rc})")
        return t

    def trace(self, n: int, k: int, seed: int = 42, first: int = 0):
        f = np.zeros(n * k, dtype=np.int32)
        kd = np.zeros(n * k, dtype=np.int32)
        lib().mfcpu_trace(self.h, seed, first, n, k, _ip(f), _ip(kd))
        return f.reshape(n, k), kd.reshape(n, k)

    def intersect(self, origins, dirs, last_hit=None, structure: int = 0, brute: bool = True):
        o = np.ascontiguousarray(origins, dtype=np.float64)
        d = np.ascontiguousarray(dirs, dtype=np.float64)
        n = len(o)
        lh = np.ascontiguousarray(last_hit if last_hit is not None else np.full(n, -1), dtype=np.int32)
        f = np.zeros(n, dtype=np.int32)
        t, u, v = np.zeros(n), np.zeros(n), np.zeros(n)
        rc = lib().mfcpu_intersect_batch(self.h, n, _dp(o), _dp(d), _ip(lh), structure, int(brute), _ip(f),
                                         _dp(t), _dp(u), _dp(v))
        if rc:
            raise RuntimeError("mfcpu_intersect_batch failed")
        return f, t, u, v

    def bvh_stats(self, origins, dirs, last_hit=None, structure: int = 0):
        o = np.ascontiguousarray(origins, dtype=np.float64)
        d = np.ascontiguousarray(dirs, dtype=np.float64)
        n = len(o)
        lh = np.ascontiguousarray(last_hit if last_hit is not None else np.full(n, -1), dtype=np.int32)
        a, b = C.c_double(), C.c_double()
        lib().mfcpu_bvh_stats(self.h, n, _dp(o), _dp(d), _ip(lh), structure, C.byref(a), C.byref(b))
        return a.value, b.value

    def area(self, f: int) -> float:
        return lib().mfcpu_facet_area(self.h, f)

    def frame(self, f: int) -> np.ndarray:
        o = np.zeros(12)
        lib().mfcpu_facet_frame(self.h, f, _dp(o))
        return o.reshape(4, 3)

    def texture_increments(self, f: int, n: int) -> np.ndarray:
        o = np.zeros(n)
        lib().mfcpu_texture_increments(self.h, f, _dp(o), n)
        return o

    @property
    def total_desorbed_molecules(self) -> float:
        return lib().mfcpu_total_desorbed_molecules(self.h)

    @property
    def final_outgassing_rate(self) -> float:
        return lib().mfcpu_final_outgassing_rate(self.h)

    @property
    def latest_moment(self) -> float:
        return lib().mfcpu_latest_moment(self.h)

    def adjoint(self, n_des, group, sticking, targets, seed=42, first=0, threads=0, max_events=0):
        group = np.ascontiguousarray(group, dtype=np.int32)
        sticking = np.ascontiguousarray(sticking, dtype=np.float64)
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        j0 = np.zeros(len(targets))
        jac = np.zeros((len(targets), len(sticking)))
        rc = lib().mfcpu_adjoint_run(self.h, seed, first, n_des, _ip(group), len(sticking), _dp(sticking),
                                     _ip(targets), len(targets), threads, _dp(j0), _dp(jac), max_events)
        if rc:
            raise RuntimeError("mfcpu_adjoint_run failed (low-flux mode required)")
        return j0, jac


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().mfcpu_philox(c, k, o)
    return list(o)


def scatter_sample(mass_ratio: float, rho: float, n: int, seed: int = 1):
    """(cos of the deflection, new speed) of n unit-speed scatters."""
    c, v = np.zeros(n), np.zeros(n)
    rc = lib().mfcpu_scatter_sample(mass_ratio, rho, n, seed, _dp(c), _dp(v))
    if rc:
        raise RuntimeError("mfcpu_scatter_sample failed")
    return c, v


def maxwell_cdf(gas_mass: float):
    v, p = np.zeros(100), np.zeros(100)
    lib().mfcpu_maxwell_cdf(gas_mass, _dp(v), _dp(p), 100)
    return v, p


def parse_user_moment(content: str, window: float):
    out = np.zeros(100000)
    n = C.c_size_t()
    lib().mfcpu_parse_user_moment(content.encode(), window, _dp(out), len(out), C.byref(n))
    return out[:n.value].copy()


def integrated_desorption(x, y, log_x, log_y, latest):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ot, oc = np.zeros(4096), np.zeros(4096)
    n = C.c_size_t()
    lib().mfcpu_integrated_desorption(_dp(x), _dp(y), len(x), int(log_x), int(log_y), latest, _dp(ot), _dp(oc),
                                      4096, C.byref(n))
    return ot[:n.value].copy(), oc[:n.value].copy()
