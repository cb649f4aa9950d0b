"""Host-side model container: the Python mirror of MolflowSimulationModel.

A Model holds vertices, facets (polygon index lists plus one mfg_facet_desc
per facet), global settings, time-dependent parameters and user moments,
and produces the flat mfg_model_view the C ABI consumes. Defaults follow
the reference XML loader (src/IO/LoaderXML.cpp:831-1126) and the facet
constructor defaults it relies on.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from ._abi import MFG_ABI_VERSION, FacetDesc, ModelView, Settings

FACET_DTYPE = np.dtype(FacetDesc)


def default_facets(n: int) -> np.ndarray:
    """n facet descriptors with the reference's defaults."""
    f = np.zeros(n, dtype=FACET_DTYPE)
    f["opacity"] = 1.0
    f["temperature"] = 293.15
    f["accomodation"] = 1.0
    f["diffuse_part"] = 1.0
    f["sojourn_freq"] = 1e13
    f["sojourn_e"] = 100.0
    f["am_theta_limit"] = 1.570796326
    for k in ("sticking_param", "opacity_param", "outgassing_param", "temperature_param"):
        f[k] = -1
    return f


def texture_cells(precise: float) -> int:
    """Texture cells along one side for a precise size (texWidth from
    texWidth_precise): the ceiling, except that a size exceeding a whole
    number by rounding noise only (relative 1e-9) counts as that number.
    MolFlow's own rule lives in its absent shared library; its stored
    results (fixture 02c: texDimY 5.000000000000002, 5 rows) show this
    behaviour."""
    import math
    n = math.ceil(precise)
    if n > 1 and precise - (n - 1) <= 1e-9 * precise:
        n -= 1
    return int(n)


def default_settings() -> Settings:
    s = Settings()
    s.gas_mass = 28.0
    s.half_life = 1.0
    s.time_window = 1e-10
    s.low_flux_cutoff = 1e-7
    s.scatter_rho = 1.0      # the scattering laws depend on b / rho only (b = rho sqrt(u))
    s.use_maxwell = 1
    s.calc_constant_flow = 1
    s.log_facet_id = -1
    return s


@dataclass
class Parameter:
    """A time-dependent parameter: (t, value) points, optional log axes."""
    name: str
    x: Sequence[float]
    y: Sequence[float]
    log_x: bool = False
    log_y: bool = False


def parse_user_moment(content: str, window: float) -> list[float]:
    """ParseUserMoment (reference src/TimeMoments.cpp:58-83): "t" or
    "begin,spacing,end"; the range loop accumulates spacing in floating
    point, exactly like the reference, so the end point may drop out."""
    parts = content.split(",")
    try:
        vals = [float(p) for p in parts]
    except ValueError:
        return []
    out: list[float] = []
    if len(vals) == 1 and vals[0] >= 0.0:
        out.append(vals[0])
    elif len(vals) == 3:
        b, s, e = vals
        if b >= 0.0 and e > b and s < (e - b) and not (s < window):
            t = b
            while t <= e:
                out.append(t)
                t += s
    return out


@dataclass
class Model:
    vertices: np.ndarray                       # (nV,3) float64, cm
    polygons: list                             # list of vertex-index sequences
    facets: np.ndarray                         # FACET_DTYPE, one per polygon
    settings: Settings = field(default_factory=default_settings)
    n_structures: int = 1
    params: list = field(default_factory=list)
    moments: list = field(default_factory=list)   # (time, window) pairs
    og_maps: dict = field(default_factory=dict)   # facet -> 2D array (h, w) Pa*m3/s
    am_maps: dict = field(default_factory=dict)   # facet -> recorded angle map counts (DES_ANGLEMAP)
    name: str = "model"

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64).reshape(-1, 3)
        if len(self.facets) != len(self.polygons):
            raise ValueError("one facet descriptor per polygon required")
        self._keep: list = []

    @property
    def n_facets(self) -> int:
        return len(self.polygons)

    def add_user_moments(self, content: str, window: float) -> None:
        for t in parse_user_moment(content, window):
            self.moments.append((t, window))

    def param_index(self, name: str) -> int:
        for i, p in enumerate(self.params):
            if p.name == name:
                return i
        raise KeyError(name)

    def facet_areas(self) -> np.ndarray:
        """Polygon areas (cm^2) from the 3D vertices (Newell's formula)."""
        out = np.empty(self.n_facets)
        by_len: dict[int, list[int]] = {}
        for i, poly in enumerate(self.polygons):
            by_len.setdefault(len(poly), []).append(i)
        for k, ids in by_len.items():
            # all polygons with k vertices at once: P[f, j] and its successor
            idx = np.asarray([self.polygons[i] for i in ids], dtype=np.int64)
            P = self.vertices[idx]
            Q = np.roll(P, -1, axis=1)
            n = np.stack([np.sum((P[..., 1] - Q[..., 1]) * (P[..., 2] + Q[..., 2]), axis=1),
                          np.sum((P[..., 2] - Q[..., 2]) * (P[..., 0] + Q[..., 0]), axis=1),
                          np.sum((P[..., 0] - Q[..., 0]) * (P[..., 1] + Q[..., 1]), axis=1)], axis=1)
            out[np.asarray(ids)] = 0.5 * np.linalg.norm(n, axis=1)
        return out

    def view(self) -> ModelView:
        """Flat C view. Arrays are kept alive on the Model until the next call."""
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a

        off = np.zeros(self.n_facets + 1, dtype=np.uint32)
        lens = np.fromiter((len(p) for p in self.polygons), dtype=np.int64, count=self.n_facets)
        off[1:] = np.cumsum(lens)
        idx = arr(np.concatenate([np.asarray(p, dtype=np.uint32) for p in self.polygons]) if self.n_facets else
                  np.zeros(0), np.uint32)
        off = arr(off, np.uint32)
        verts = arr(self.vertices.reshape(-1), np.float64)
        facets = arr(self.facets, FACET_DTYPE)
        poff = np.zeros(len(self.params) + 1, dtype=np.uint32)
        px, py, lx, ly = [], [], [], []
        for i, p in enumerate(self.params):
            poff[i + 1] = poff[i] + len(p.x)
            px.extend(p.x); py.extend(p.y); lx.append(int(p.log_x)); ly.append(int(p.log_y))
        poff = arr(poff, np.uint32)
        px = arr(px, np.float64); py = arr(py, np.float64)
        lx = arr(lx, np.uint8); ly = arr(ly, np.uint8)
        mt = arr([m[0] for m in self.moments], np.float64)
        mw = arr([m[1] for m in self.moments], np.float64)
        ogo = np.zeros(self.n_facets + 1, dtype=np.uint64)
        vals = []
        for f in range(self.n_facets):
            m = self.og_maps.get(f)
            n = 0 if m is None else int(np.asarray(m).size)
            ogo[f + 1] = ogo[f] + n
            if m is not None:
                vals.append(np.asarray(m, dtype=np.float64).reshape(-1))
        ogo = arr(ogo, np.uint64)
        ogv = arr(np.concatenate(vals) if vals else np.zeros(1), np.float64)
        amo = np.zeros(self.n_facets + 1, dtype=np.uint64)
        amv = []
        for f in range(self.n_facets):
            a = self.am_maps.get(f)
            amo[f + 1] = amo[f] + (0 if a is None else int(np.asarray(a).size))
            if a is not None:
                amv.append(np.asarray(a, dtype=np.uint64).reshape(-1))
        amo = arr(amo, np.uint64)
        amv = arr(np.concatenate(amv) if amv else np.zeros(1), np.uint64)
        settings = Settings()
        C.memmove(C.byref(settings), C.byref(self.settings), C.sizeof(Settings))
        keep.append(settings)

        def ptr(a, ct):
            return a.ctypes.data_as(C.POINTER(ct))

        v = ModelView()
        v.struct_size = C.sizeof(ModelView)
        v.abi_version = MFG_ABI_VERSION
        v.n_vertices = len(self.vertices)
        v.n_facets = self.n_facets
        v.n_structures = self.n_structures
        v.n_params = len(self.params)
        v.n_moments = len(self.moments)
        v.vertices = ptr(verts, C.c_double)
        v.facet_vertex_offset = ptr(off, C.c_uint32)
        v.facet_vertex_index = ptr(idx, C.c_uint32)
        v.facets = facets.ctypes.data_as(C.POINTER(FacetDesc))
        v.settings = C.pointer(settings)
        v.param_offset = ptr(poff, C.c_uint32)
        v.param_x = ptr(px, C.c_double)
        v.param_y = ptr(py, C.c_double)
        v.param_log_x = ptr(lx, C.c_uint8)
        v.param_log_y = ptr(ly, C.c_uint8)
        v.moment_time = ptr(mt, C.c_double)
        v.moment_window = ptr(mw, C.c_double)
        if self.og_maps:
            v.og_map_offset = ptr(ogo, C.c_uint64)
            v.og_map_values = ptr(ogv, C.c_double)
        if self.am_maps:
            v.am_pdf_offset = ptr(amo, C.c_uint64)
            v.am_pdf = ptr(amv, C.c_uint64)
        self._keep = keep
        return v
