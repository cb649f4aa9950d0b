"""GPU simulation unit over the C ABI: the Python mirror of MolflowSimulation
(reference src/Simulation/MolflowSimulation.h:22-45).

    sim = GpuSimulation()                 # ConstructParticleTracers role
    sim.load(model)                       # LoadSimulation + PrepareToRun
    sim.rebuild_accel()                   # RebuildAccelStructure
    sim.reset(seed=42)                    # ResetSimulation
    res = sim.run(n_des)                  # desorb n_des particles to the end
    t = sim.read()                        # UpdateMCHits into a fresh state
    sim.take(t)                           # UpdateHitsAndLog: add, then clear
    problems = sanity_check(model)        # SanityCheckModel (no device needed)

Every method raises MolflowGpuError with the library's message on failure;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _abi as A
from .tallies import OFFSET_KINDS, Tallies


class MolflowGpuError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            _lib = A.load_library()
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class GpuSimulation:
    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.mfg_create(C.byref(h), device)
        if rc:
            raise MolflowGpuError(f"mfg_create failed ({rc}): no usable HIP device {device}")
        self.h = h
        self.model = None
        self.models = None
        self._views = []
        self.generation = 0      # counts runs; the adjoint binding checks it

    def close(self):
        if getattr(self, "h", None):
            self._lib.mfg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            msg = self._lib.mfg_last_error(self.h)
            raise MolflowGpuError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- simulation-unit API
    def load(self, model) -> None:
        v = model.view()
        self._views = [v]
        self._check(self._lib.mfg_load_model(self.h, C.byref(v)), "mfg_load_model")
        self.model, self.models = model, None

    def load_variants(self, models) -> None:
        views = [m.view() for m in models]
        arr = (C.POINTER(A.ModelView) * len(views))(*[C.pointer(v) for v in views])
        self._views = views
        self._check(self._lib.mfg_load_variants(self.h, arr, len(views)), "mfg_load_variants")
        self.models, self.model = list(models), None

    def rebuild_accel(self, width: int = 8, leaf_size: int = 2) -> None:
        self._check(self._lib.mfg_build_accel(self.h, width, leaf_size), "mfg_build_accel")

    def reset(self, seed: int = 42, particle_offset: int = 0) -> None:
        self.generation += 1
        self._check(self._lib.mfg_reset(self.h, seed, particle_offset), "mfg_reset")

    def run(self, n_des: int, max_events: int = 0) -> A.StepResult:
        self.generation += 1
        r = A.StepResult()
        self._check(self._lib.mfg_run(self.h, int(n_des), int(max_events), C.byref(r)), "mfg_run")
        return r

    def run_variants(self, n_des, seeds, max_events: int = 0) -> A.StepResult:
        self.generation += 1
        n = np.ascontiguousarray(n_des, dtype=np.uint64)
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        r = A.StepResult()
        self._check(self._lib.mfg_run_variants(self.h, n.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               s.ctypes.data_as(C.POINTER(C.c_uint64)), int(max_events),
                                               C.byref(r)), "mfg_run_variants")
        return r

    def request_stop(self) -> None:
        self._check(self._lib.mfg_request_stop(self.h), "mfg_request_stop")

    def set_time_limit(self, seconds: float) -> None:
        """Wall-clock limit of one trace launch; a run it ends reports stopped = 2."""
        self._check(self._lib.mfg_set_time_limit(self.h, float(seconds)), "mfg_set_time_limit")

    def set_launch(self, block: int = 256, blocks_per_cu: int = 0) -> None:
        """Threads per block (a multiple of 64 up to 256, or 768 / 1024 for
        one block per CU at 3 / 4 waves per SIMD) and blocks per CU (0 =
        fill each CU)."""
        self._check(self._lib.mfg_set_launch(self.h, block, blocks_per_cu), "mfg_set_launch")

    def set_carry(self, on: bool = True) -> None:
        """Keep particles in flight between runs (see mfg_set_carry)."""
        self._check(self._lib.mfg_set_carry(self.h, int(on)), "mfg_set_carry")

    def in_flight(self) -> int:
        n = C.c_uint64()
        self._check(self._lib.mfg_in_flight(self.h, C.byref(n)), "mfg_in_flight")
        return n.value

    def drain(self, max_events: int = 0) -> A.StepResult:
        """Trace every parked particle to its end (carry is switched off)."""
        self.set_carry(False)
        return self.run(0, max_events=max_events) if self.models is None else \
            self.run_variants(np.zeros(len(self.models), dtype=np.uint64), np.zeros(len(self.models), dtype=np.uint64),
                              max_events=max_events)

    def set_kernel_variant(self, waves: int = 0) -> None:
        self._check(self._lib.mfg_set_kernel_variant(self.h, waves), "mfg_set_kernel_variant")

    def autotune(self, n_probe: int = 100_000) -> A.TuneResult:
        r = A.TuneResult()
        self._check(self._lib.mfg_autotune(self.h, int(n_probe), C.byref(r)), "mfg_autotune")
        return r

    def enable_graph(self, on: bool = True) -> None:
        self._check(self._lib.mfg_enable_graph(self.h, int(on)), "mfg_enable_graph")

    def layout(self, variant: int | None = None) -> A.TallyLayout:
        L = A.TallyLayout()
        if variant is None:
            self._check(self._lib.mfg_get_layout(self.h, C.byref(L)), "mfg_get_layout")
        else:
            self._check(self._lib.mfg_get_variant_layout(self.h, variant, C.byref(L)), "mfg_get_variant_layout")
        return L

    def offsets(self) -> dict:
        n = self.layout().n_facets
        out = {}
        for name, k in OFFSET_KINDS.items():
            a = np.zeros(n, dtype=np.int64)
            self._check(self._lib.mfg_get_facet_offsets(self.h, k, a.ctypes.data_as(C.POINTER(C.c_int64)), n),
                        "mfg_get_facet_offsets")
            out[name] = a
        return out

    def read(self, tallies: Tallies | None = None, variant: int | None = None) -> Tallies:
        if tallies is None:
            tallies = Tallies(self.layout(variant), self.offsets() if variant is None else None)
        v = tallies.view()
        if variant is None:
            self._check(self._lib.mfg_read_tallies(self.h, C.byref(v)), "mfg_read_tallies")
        else:
            self._check(self._lib.mfg_read_variant_tallies(self.h, variant, C.byref(v)), "mfg_read_variant_tallies")
        return tallies

    def take(self, tallies: Tallies | None = None) -> Tallies:
        """Add the device tallies into `tallies` and clear them on the device,
        keeping parked particles (a periodic report, UpdateHitsAndLog)."""
        if tallies is None:
            tallies = Tallies(self.layout(), self.offsets())
        v = tallies.view()
        self._check(self._lib.mfg_take_tallies(self.h, C.byref(v)), "mfg_take_tallies")
        return tallies

    def clear_tallies(self) -> None:
        self._check(self._lib.mfg_clear_tallies(self.h), "mfg_clear_tallies")

    def reinit_particle_log(self) -> None:
        self._check(self._lib.mfg_reinit_particle_log(self.h), "mfg_reinit_particle_log")

    def hits_size(self) -> int:
        b = C.c_uint64()
        self._check(self._lib.mfg_hits_size(self.h, C.byref(b)), "mfg_hits_size")
        return b.value

    def intersect(self, origins, dirs, last_hit=None, structure: int = 0):
        o = np.ascontiguousarray(origins, dtype=np.float64)
        d = np.ascontiguousarray(dirs, dtype=np.float64)
        n = len(o)
        lh = None if last_hit is None else np.ascontiguousarray(last_hit, dtype=np.int32)
        f = np.zeros(n, dtype=np.int32)
        t, u, v = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self._lib.mfg_intersect_batch(self.h, n, _dp(o), _dp(d), None if lh is None else _ip(lh),
                                                  structure, _ip(f), _dp(t), _dp(u), _dp(v)),
                    "mfg_intersect_batch")
        return f, t, u, v

    def intersect_static(self, origins, dirs, last_hit=None, structure: int = 0):
        """intersect() through the entry point that allocates no device
        memory (at most STATIC_MAX_RAYS rays per call)."""
        o = np.ascontiguousarray(origins, dtype=np.float64)
        d = np.ascontiguousarray(dirs, dtype=np.float64)
        n = len(o)
        lh = None if last_hit is None else np.ascontiguousarray(last_hit, dtype=np.int32)
        f = np.zeros(n, dtype=np.int32)
        t, u, v = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self._lib.mfg_intersect_batch_static(self.h, n, _dp(o), _dp(d), None if lh is None else _ip(lh),
                                                         structure, _ip(f), _dp(t), _dp(u), _dp(v)),
                    "mfg_intersect_batch_static")
        return f, t, u, v

    def set_trace(self, n: int, k: int) -> None:
        self._trace_shape = (n, k)
        self._check(self._lib.mfg_set_trace(self.h, n, k), "mfg_set_trace")

    def get_trace(self):
        n, k = self._trace_shape
        f = np.zeros(n * k, dtype=np.int32)
        kd = np.zeros(n * k, dtype=np.int32)
        self._check(self._lib.mfg_get_trace(self.h, _ip(f), _ip(kd), n * k), "mfg_get_trace")
        return f.reshape(n, k), kd.reshape(n, k)

    # ---- adjoint
    def adjoint_setup(self, group_of_facet, targets) -> None:
        g = np.ascontiguousarray(group_of_facet, dtype=np.int32)
        t = np.ascontiguousarray(targets, dtype=np.int32)
        self._adj_shape = (len(t), int(g.max()) + 1)
        self._check(self._lib.mfg_adjoint_setup(self.h, _ip(g), self._adj_shape[1], _ip(t), len(t)),
                    "mfg_adjoint_setup")

    def adjoint_set_sticking(self, s) -> None:
        s = np.ascontiguousarray(s, dtype=np.float64)
        self._check(self._lib.mfg_adjoint_set_sticking(self.h, _dp(s), len(s)), "mfg_adjoint_set_sticking")

    def adjoint_read(self):
        nt, ng = self._adj_shape
        j0, jac = np.zeros(nt), np.zeros((nt, ng))
        self._check(self._lib.mfg_adjoint_read(self.h, _dp(j0), _dp(jac)), "mfg_adjoint_read")
        return j0, jac

    def adjoint_backward(self, upstream) -> np.ndarray:
        nt, ng = self._adj_shape
        up = np.ascontiguousarray(upstream, dtype=np.float64)
        g = np.zeros(ng)
        self._check(self._lib.mfg_adjoint_backward(self.h, _dp(up), _dp(g)), "mfg_adjoint_backward")
        return g


def sanity_check(model) -> list[str]:
    """SanityCheckModel (reference src/Simulation/MolflowSimulation.cpp:57-71):
    the problems that would stop `model` from loading, plus the reference's
    own warnings; an empty list means the model loads. Needs no device."""
    v = model.view()
    buf = C.create_string_buffer(65536)
    n = C.c_uint32()
    rc = lib().mfg_sanity_check(C.byref(v), buf, len(buf), C.byref(n))
    if rc:
        raise MolflowGpuError(f"mfg_sanity_check failed ({rc})")
    return [line for line in buf.value.decode().split("\n") if line][: n.value]


def maxwell_cdf(gas_mass: float):
    v, p = np.zeros(100), np.zeros(100)
    rc = lib().mfg_maxwell_cdf(gas_mass, _dp(v), _dp(p), 100)
    if rc:
        raise MolflowGpuError("mfg_maxwell_cdf failed")
    return v, p


def parse_user_moment(content: str, window: float):
    out = np.zeros(100000)
    n = C.c_size_t()
    rc = lib().mfg_parse_user_moment(content.encode(), window, _dp(out), len(out), C.byref(n))
    if rc:
        raise MolflowGpuError("mfg_parse_user_moment failed")
    return out[:n.value].copy()


def integrated_desorption(x, y, log_x, log_y, latest):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ot, oc = np.zeros(4096), np.zeros(4096)
    n = C.c_size_t()
    rc = lib().mfg_integrated_desorption(_dp(x), _dp(y), len(x), int(log_x), int(log_y), latest, _dp(ot), _dp(oc),
                                         4096, C.byref(n))
    if rc:
        raise MolflowGpuError("mfg_integrated_desorption failed")
    return ot[:n.value].copy(), oc[:n.value].copy()
