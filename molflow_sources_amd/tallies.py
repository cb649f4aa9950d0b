"""Host tally buffers laid out like GlobalSimuState
(reference src/Simulation/MolflowSimGeom.h:56-154), allocated from an
mfg_tally_layout and exposed to the C ABI as an mfg_tally_view."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import TallyLayout, TallyView, _np_dtypes

FACET_DT, TEX_DT, PROF_DT, DIR_DT, LOG_DT, GLOBAL_DT = _np_dtypes()


class Tallies:
    def __init__(self, layout: TallyLayout, offsets: dict | None = None):
        L = layout
        self.layout = L
        nF, S = L.n_facets, L.n_slots
        self.n_facets, self.n_slots = nF, S
        self.glob = np.zeros(1, dtype=GLOBAL_DT)
        self.facets = np.zeros((S, nF), dtype=FACET_DT)
        self.forces = np.zeros((S, nF, 9))
        self.textures = np.zeros((S, L.n_tex_cells), dtype=TEX_DT)
        self.directions = np.zeros((S, L.n_dir_cells), dtype=DIR_DT)
        self.profiles = np.zeros((S, L.n_profile_slices), dtype=PROF_DT)
        self.angle_maps = np.zeros(L.n_angle_cells, dtype=np.uint64)
        self.global_hist = np.zeros((S, L.n_global_hist_bins))
        self.facet_hist = np.zeros((S, L.n_facet_hist_bins))
        self.log = np.zeros(max(1, L.log_capacity), dtype=LOG_DT)
        self.log_count = np.zeros(1, dtype=np.uint64)
        self.leaks = np.zeros((max(1, L.leak_capacity), 6))
        self.leak_count = np.zeros(1, dtype=np.uint64)
        self.offsets = offsets or {}

    def view(self) -> TallyView:
        v = TallyView()
        v.struct_size = C.sizeof(TallyView)
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        v.global_ = p(self.glob)
        v.facets = p(self.facets)
        v.forces = p(self.forces)
        v.textures = p(self.textures)
        v.directions = p(self.directions)
        v.profiles = p(self.profiles)
        v.angle_maps = p(self.angle_maps)
        v.global_hist = p(self.global_hist)
        v.facet_hist = p(self.facet_hist)
        v.log = p(self.log)
        v.log_count = p(self.log_count)
        v.leak_pos_dir = p(self.leaks)
        v.leak_count = p(self.leak_count)
        return v

    # convenient accessors -------------------------------------------------
    @property
    def hits(self) -> int:
        return int(self.glob["nb_mc_hit"][0])

    @property
    def desorbed(self) -> int:
        return int(self.glob["nb_desorbed"][0])

    def facet_field(self, name: str, slot: int = 0) -> np.ndarray:
        return self.facets[slot][name].astype(np.float64)

    def texture_of(self, facet: int, w: int, h: int, slot: int = 0) -> np.ndarray:
        off = int(self.offsets["texture"][facet])
        return self.textures[slot][off:off + w * h].reshape(h, w)

    def profile_of(self, facet: int, slot: int = 0) -> np.ndarray:
        off = int(self.offsets["profile"][facet])
        return self.profiles[slot][off:off + 100]

    def angle_map_of(self, facet: int, n_theta: int, n_phi: int) -> np.ndarray:
        off = int(self.offsets["anglemap"][facet])
        return self.angle_maps[off:off + n_theta * n_phi].reshape(n_theta, n_phi)

    def __iadd__(self, other: "Tallies") -> "Tallies":
        for name in ("facets", "textures", "profiles"):
            a, b = getattr(self, name), getattr(other, name)
            for fld in a.dtype.names:
                a[fld] += b[fld]
        self.directions["dir"] += other.directions["dir"]
        self.directions["count"] += other.directions["count"]
        for fld in self.glob.dtype.names:
            self.glob[fld] += other.glob[fld]
        self.forces += other.forces
        self.angle_maps += other.angle_maps
        self.global_hist += other.global_hist
        self.facet_hist += other.facet_hist
        return self


OFFSET_KINDS = {"texture": 0, "direction": 1, "profile": 2, "anglemap": 3, "facet_hist": 4}
