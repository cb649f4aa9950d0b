"""Build Models from the converted reference fixtures (tests/golden/*.json.gz)."""
from __future__ import annotations

import gzip
import json
import os

import numpy as np

from molflow_sources_amd import _abi as A
from molflow_sources_amd.model import Model, Parameter, default_facets, default_settings

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name: str) -> dict:
    with gzip.open(os.path.join(GOLDEN, name + ".json.gz"), "rt", encoding="utf-8") as fh:
        return json.load(fh)


def model_from_fixture(name: str):
    d = load_json(name)
    polys = d["polygons"]
    F = default_facets(len(polys))
    for k, vals in d["facets"].items():
        if k in F.dtype.names:
            F[k] = np.asarray(vals, dtype=F.dtype[k])
    s = default_settings()
    for k, v in d["settings"].items():
        if isinstance(v, list):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, type(getattr(s, k))(v))
    m = Model(np.asarray(d["vertices"]), polys, F, settings=s, n_structures=d["n_structures"], name=name)
    m.params = [Parameter(p["name"], p["x"], p["y"], p["log_x"], p["log_y"]) for p in d["params"]]
    for content, window in d["user_moments"]:
        m.add_user_moments(content, window)
    # recorded incident-angle maps that DES_ANGLEMAP facets desorb from; the
    # stored map's shape replaces the facet's angle-map parameters, as the
    # reference loader does (LoaderXML.cpp:1043-1049)
    for f, am in d.get("angle_maps_in", {}).items():
        f = int(f)
        m.am_maps[f] = np.asarray(am["map"], dtype=np.uint64)
        m.facets["am_phi_width"][f] = am["phi_width"]
        m.facets["am_theta_limit"][f] = am["theta_limit"]
        m.facets["am_theta_lower_res"][f] = am["lower"]
        m.facets["am_theta_higher_res"][f] = am["higher"]
    return m, d


def stored_facets(d: dict, slot: int = 0) -> dict:
    """Stored per-facet counters of one moment slot, named like mfg_facet_counters."""
    r = d["results"]["facets"]
    names = {"nbHit": "nb_mc_hit", "nbHitEquiv": "nb_hit_equiv", "nbDes": "nb_desorbed",
             "nbAbsEquiv": "nb_abs_equiv", "sum_v_ort": "sum_v_ort", "sum_1_per_v": "sum_1_per_ort_velocity",
             "sum_v": "sum_1_per_velocity"}
    return {names[k]: np.asarray(v[slot]) for k, v in r.items()}


class StoredState:
    """Stored results shaped like Tallies, with the per-facet accessors of
    molflow_sources_amd.compare.TallyState (for compare_states): profiles
    of every stored moment, textures and histograms of the steady state."""

    def __init__(self, d: dict):
        self._r = d["results"]
        from molflow_sources_amd.tallies import FACET_DT, GLOBAL_DT
        r = d["results"]
        g = r["global"]
        self.glob = np.zeros(1, dtype=GLOBAL_DT)
        self.glob["nb_mc_hit"] = g["totalHit"]
        self.glob["nb_hit_equiv"] = g.get("totalHitEquiv", g["totalHit"])
        self.glob["nb_desorbed"] = g["totalDes"]
        self.glob["nb_abs_equiv"] = g["totalAbsEquiv"]
        self.glob["nb_leak_total"] = g.get("totalLeak", 0)
        nS, nF = r["n_slots"], len(d["polygons"])
        self.facets = np.zeros((nS, nF), dtype=FACET_DT)
        for s in range(nS):
            for k, v in stored_facets(d, s).items():
                self.facets[s][k] = v

    def profile(self, f: int, m: int):
        p = self._r["profiles"].get(str(f))
        if p is None or m >= len(p):
            return None
        a = np.asarray(p[m], dtype=np.float64)
        return a[:, 0], a[:, 1], a[:, 2]

    def texture(self, f: int, m: int):
        t = self._r["textures"].get(str(f))
        if t is None or m != 0:
            return None
        return (np.asarray(t["count"], dtype=np.float64), np.asarray(t["sum_v_ort"], dtype=np.float64),
                np.asarray(t["sum_1_per_v"], dtype=np.float64))

    def direction(self, f: int, m: int):
        return None

    def facet_hist(self, f: int, m: int):
        h = self._r["facet_hist"].get(str(f))
        if h is None or m != 0:
            return None
        return tuple(np.asarray(h.get(k, []), dtype=np.float64) for k in ("Bounces", "Distance", "Time"))

    def global_hist(self, m: int):
        h = self._r["global_hist"]
        if not h or m != 0:
            return None
        return tuple(np.asarray(h.get(k, []), dtype=np.float64) for k in ("Bounces", "Distance", "Time"))


ALL_FIXTURES = sorted(n[:-8] for n in os.listdir(GOLDEN) if n.endswith(".json.gz"))
A = A
