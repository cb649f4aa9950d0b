"""Result comparison: the reference's parity checker and sigma-based tests.

compare_states restates GlobalSimuState::Compare
(reference src/Simulation/MolflowSimGeom.cpp:558-1148) in full:

  global   abs/des and hits/des ratios (threshold `glob`, 0.009 in the
           reference's ResultsOkay, 0.006 in ResultsWrong), the global
           velocity sums (:622-636; the tracer never fills them, so they
           compare 0 with 0 unless a state holds them), and the global
           bounce / distance / time histograms of every moment, each bin
           over the global MC hits (`loc`, bins below 80^2 skipped, :639-684)
  facets   per facet and moment slot, skipped below 40000 MC hits on
           either side, with the local threshold max(loc, 1/sqrt(nMC)):
           hit equivalents and sum_v_ort over the global hit equivalents,
           the 1/velocity sums over (hits + abs + des) with a 40x looser
           threshold, abs/hits and des/hits (:703-845)
  fine     profiles smoothed by [1,4,6,4,1]/16 (the reference smooths the
           counts also where it means the velocity sums, :852-908, and so
           does this), texture cells, direction cells and the facet
           histograms (:957-1124)

It returns (global errors, facet errors, fine errors, messages); the
reference's tests fail a run on global or facet errors only. IsEqual,
absent from the checkout, is taken as |a-b| <= tol * max(|a|,|b|).

The sigma-based helpers add what the reference lacks: a binomial z-score
for transmission-type ratios, a multinomial chi-square over per-facet
absorption counts (each particle is absorbed once), and batch-means
standard errors for correlated hit tallies.
"""
from __future__ import annotations

import math

import numpy as np

VELOCITY_FACTOR = 40.0        # velocityThresholdFactor (MolflowSimGeom.cpp:562)


def is_equal(a: float, b: float, tol: float) -> bool:
    return abs(a - b) <= tol * max(abs(a), abs(b)) or (a == b)


def _too_few(a: float, b: float, level: float = 80.0) -> bool:
    """The reference's "sample size not large enough" test: sqrt(max(1,
    min(a, b))) < level."""
    return math.sqrt(max(1.0, min(float(a), float(b)))) < level


class TallyState:
    """A Tallies object (GPU or oracle run) seen the way Compare reads a
    GlobalSimuState: per-facet profiles, textures, directions and
    histograms sliced out of the flat arrays using the model's sizes."""

    def __init__(self, tallies, model):
        self.t = tallies
        self.glob = tallies.glob
        self.facets = tallies.facets
        from .model import texture_cells
        F = model.facets
        self.tex_w = np.array([texture_cells(x) for x in F["tex_width_precise"]], dtype=np.int64)
        self.tex_h = np.array([texture_cells(x) for x in F["tex_height_precise"]], dtype=np.int64)
        s = model.settings

        def hist_sizes(rb, bb, bm, rd, db, dm, rt, tb, tm):
            nb = int(bm // bb + 2) if rb and bb else 0
            nd = int(int(dm / db) + 1) if rd and db > 0 else 0
            nt = int(int(tm / tb) + 1) if rt and tb > 0 else 0
            return nb, nd, nt

        self.gh_sizes = hist_sizes(s.gh_record_bounce, s.gh_bounce_binsize, s.gh_bounce_max, s.gh_record_distance,
                                   s.gh_distance_binsize, s.gh_distance_max, s.gh_record_time, s.gh_time_binsize,
                                   s.gh_time_max)
        self.fh_sizes = [hist_sizes(f["fh_record_bounce"], f["fh_bounce_binsize"], f["fh_bounce_max"],
                                    f["fh_record_distance"], f["fh_distance_binsize"], f["fh_distance_max"],
                                    f["fh_record_time"], f["fh_time_binsize"], f["fh_time_max"]) for f in F]

    @property
    def n_slots(self) -> int:
        return self.facets.shape[0]

    def profile(self, f: int, m: int):
        off = int(self.t.offsets["profile"][f])
        if off < 0:
            return None
        p = self.t.profiles[m][off:off + 100]
        return p["count_equiv"], p["sum_v_ort"], p["sum_1_per_ort_velocity"]

    def texture(self, f: int, m: int):
        off = int(self.t.offsets["texture"][f])
        if off < 0:
            return None
        n = int(self.tex_w[f] * self.tex_h[f])
        c = self.t.textures[m][off:off + n]
        return c["count_equiv"], c["sum_v_ort_per_area"], c["sum_1_per_ort_velocity"]

    def direction(self, f: int, m: int):
        off = int(self.t.offsets["direction"][f])
        if off < 0:
            return None
        n = int(self.tex_w[f] * self.tex_h[f])
        c = self.t.directions[m][off:off + n]
        return c["count"].astype(np.float64), c["dir"]

    def facet_hist(self, f: int, m: int):
        off = int(self.t.offsets["facet_hist"][f])
        if off < 0:
            return None
        nb, nd, nt = self.fh_sizes[f]
        h = self.t.facet_hist[m][off:off + nb + nd + nt]
        return h[:nb], h[nb:nb + nd], h[nb + nd:]

    def global_hist(self, m: int):
        nb, nd, nt = self.gh_sizes
        if nb + nd + nt == 0:
            return None
        h = self.t.global_hist[m]
        return h[:nb], h[nb:nb + nd], h[nb + nd:nb + nd + nt]


def _get(state, name, *args):
    fn = getattr(state, name, None)
    return fn(*args) if fn is not None else None


def compare_states(lhs, rhs, glob: float = 0.009, loc: float = 0.07, moments: bool = True, counts=None):
    """GlobalSimuState::Compare of two states. lhs/rhs expose .glob (global
    counters), .facets[slot][facet] counters and, optionally, the accessors
    of TallyState (profile, texture, direction, facet_hist, global_hist;
    None where a state holds no such record). Returns (n_global_errors,
    n_facet_errors, n_fine_errors, messages); `counts`, a dict, receives
    how many facet and fine values were actually compared."""
    msgs = []
    g_err = f_err = fine = 0
    n_facet = n_fine = 0
    L, R = lhs.glob[0], rhs.glob[0]
    if L["nb_desorbed"] == 0 or R["nb_desorbed"] == 0:
        return 1, -1, -1, ["[Global][desorp] a state has no recorded desorptions"]
    a1, a2 = L["nb_abs_equiv"] / L["nb_desorbed"], R["nb_abs_equiv"] / R["nb_desorbed"]
    if not is_equal(a1, a2, glob):
        g_err += 1; msgs.append(f"[Global][absRatio abs/des] {a1} vs {a2}")
    h1, h2 = L["nb_mc_hit"] / L["nb_desorbed"], R["nb_mc_hit"] / R["nb_desorbed"]
    if not is_equal(h1, h2, glob):
        g_err += 1; msgs.append(f"[Global][hits/des] {h1} vs {h2}")
    for fld in ("sum_v_ort", "sum_1_per_velocity", "sum_1_per_ort_velocity"):
        x = float(L[fld]) if fld in (L.dtype.names or ()) else 0.0
        y = float(R[fld]) if fld in (R.dtype.names or ()) else 0.0
        if not is_equal(x, y, glob):
            g_err += 1; msgs.append(f"[Global][{fld}] {x} vs {y}")
    n_slots = min(lhs.facets.shape[0], rhs.facets.shape[0]) if moments else 1
    # global histograms, per moment, bins over the global MC hits
    for m in range(n_slots):
        hl, hr = _get(lhs, "global_hist", m), _get(rhs, "global_hist", m)
        if hl is None or hr is None:
            continue
        for kind, xl, xr in zip(("Bounces", "Dist", "Time"), hl, hr):
            for i in range(min(len(xl), len(xr))):
                if _too_few(xl[i], xr[i]):
                    continue
                x, y = xl[i] / L["nb_mc_hit"], xr[i] / R["nb_mc_hit"]
                if not is_equal(x, y, loc):
                    g_err += 1; msgs.append(f"[Global][Hist][{kind}][moment {m}][{i}] {x} vs {y}")
    sc, scr = 1.0 / L["nb_hit_equiv"], 1.0 / R["nb_hit_equiv"]
    fs = 1.0 / (L["nb_hit_equiv"] + L["nb_abs_equiv"] + L["nb_desorbed"])
    fsr = 1.0 / (R["nb_hit_equiv"] + R["nb_abs_equiv"] + R["nb_desorbed"])
    if lhs.facets.shape[1] != rhs.facets.shape[1]:
        f_err += 1; msgs.append("[FacetStates] different number of facets")
    for f in range(min(lhs.facets.shape[1], rhs.facets.shape[1])):
        for m in range(n_slots):
            cl, cr = lhs.facets[m][f], rhs.facets[m][f]
            if cl["nb_mc_hit"] < 40000 or cr["nb_mc_hit"] < 40000:
                continue
            n_facet += 1
            lt = max(loc, 1.0 / math.sqrt(min(cl["nb_mc_hit"], cr["nb_mc_hit"])))
            shd = cl["nb_hit_equiv"] + float(cl["nb_desorbed"])
            shd_r = cr["nb_hit_equiv"] + float(cr["nb_desorbed"])
            if not _too_few(cl["nb_hit_equiv"], cr["nb_hit_equiv"]):
                pairs = [("facet_hits / global_hits", cl["nb_hit_equiv"] * sc, cr["nb_hit_equiv"] * scr, lt),
                         ("sum_v_ort", cl["sum_v_ort"] * sc, cr["sum_v_ort"] * scr, lt),
                         ("sum_1_per_velocity", cl["sum_1_per_velocity"] * fs, cr["sum_1_per_velocity"] * fsr,
                          lt * VELOCITY_FACTOR),
                         ("sum_1_per_ort_velocity", cl["sum_1_per_ort_velocity"] * fs,
                          cr["sum_1_per_ort_velocity"] * fsr, lt * VELOCITY_FACTOR)]
                for name, x, y, tol in pairs:
                    if not is_equal(x, y, tol):
                        f_err += 1; msgs.append(f"[Facet][{f}][{name}][moment {m}] {x} vs {y}")
            if not _too_few(cl["nb_abs_equiv"], cr["nb_abs_equiv"]):
                x, y = cl["nb_abs_equiv"] / cl["nb_mc_hit"], cr["nb_abs_equiv"] / cr["nb_mc_hit"]
                if not is_equal(x, y, lt):
                    f_err += 1; msgs.append(f"[Facet][{f}][abs/hits ratio][moment {m}] {x} vs {y}")
            if not _too_few(cl["nb_desorbed"], cr["nb_desorbed"]):
                x = float(cl["nb_desorbed"]) / float(cl["nb_mc_hit"])
                y = float(cr["nb_desorbed"]) / float(cr["nb_mc_hit"])
                if not is_equal(x, y, lt):
                    f_err += 1; msgs.append(f"[Facet][{f}][desRatio][moment {m}] {x} vs {y}")
            # profiles: [1,4,6,4,1]/16 smoothing of the counts, which the
            # reference also uses in place of both velocity sums
            pl, pr = _get(lhs, "profile", f, m), _get(rhs, "profile", f, m)
            if pl is not None and pr is not None:
                kern = np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0
                sl = np.convolve(np.asarray(pl[0], dtype=np.float64), kern, mode="valid")
                sr = np.convolve(np.asarray(pr[0], dtype=np.float64), kern, mode="valid")
                for i in range(min(len(sl), len(sr))):
                    if max(1.0, min(sl[i], sr[i])) < 100:
                        continue
                    n_fine += 3
                    if not is_equal(sl[i] / shd, sr[i] / shd_r, lt):
                        fine += 1; msgs.append(f"[Facet][{f}][Profile][{i + 2}][countEquiv][moment {m}]")
                    for name in ("sum_1_per_ort_velocity", "sum_v_ort"):
                        if not is_equal(sl[i] * sc, sr[i] * scr, lt * VELOCITY_FACTOR):
                            fine += 1; msgs.append(f"[Facet][{f}][Profile][{i + 2}][{name}][moment {m}]")
            tl, tr = _get(lhs, "texture", f, m), _get(rhs, "texture", f, m)
            if tl is not None and tr is not None:
                for i in range(min(len(tl[0]), len(tr[0]))):
                    if max(1.0, min(tl[0][i], tr[0][i])) < 640:
                        continue
                    n_fine += 3
                    if not is_equal(tl[0][i] / shd, tr[0][i] / shd_r, lt):
                        fine += 1; msgs.append(f"[Facet][{f}][Texture][{i}][countEquiv][moment {m}]")
                    if not is_equal(tl[2][i] * fs, tr[2][i] * fsr, lt * VELOCITY_FACTOR):
                        fine += 1; msgs.append(f"[Facet][{f}][Texture][{i}][sum_1_per_ort_velocity][moment {m}]")
                    if not is_equal(tl[1][i] * sc, tr[1][i] * scr, lt * VELOCITY_FACTOR):
                        fine += 1; msgs.append(f"[Facet][{f}][Texture][{i}][sum_v_ort_per_area][moment {m}]")
            dl, dr = _get(lhs, "direction", f, m), _get(rhs, "direction", f, m)
            if dl is not None and dr is not None:
                for i in range(min(len(dl[0]), len(dr[0]))):
                    if _too_few(dl[0][i], dr[0][i]):
                        continue
                    n_fine += 4
                    if not is_equal(dl[0][i], dr[0][i], lt):
                        fine += 1; msgs.append(f"[Facet][{f}][dirs][{i}][count][moment {m}]")
                    for k, axis in enumerate("xyz"):
                        if not is_equal(dl[1][i][k], dr[1][i][k], lt):
                            fine += 1; msgs.append(f"[Facet][{f}][dirs][{i}][dir.{axis}][moment {m}]")
            hl, hr = _get(lhs, "facet_hist", f, m), _get(rhs, "facet_hist", f, m)
            if hl is not None and hr is not None:
                for kind, xl, xr in zip(("Bounces", "Dist", "Time"), hl, hr):
                    for i in range(min(len(xl), len(xr))):
                        if _too_few(xl[i], xr[i]):
                            continue
                        n_fine += 1
                        if not is_equal(xl[i] / cl["nb_mc_hit"], xr[i] / cr["nb_mc_hit"], lt):
                            fine += 1; msgs.append(f"[Facet][{f}][Hist][{kind}][{i}][moment {m}]")
    if counts is not None:
        counts["facet_checks"] = n_facet
        counts["fine_checks"] = n_fine
    return g_err, f_err, fine, msgs


def results_wrong_detection(run_short, stored, n_runs: int = 15, glob: float = 0.006, loc: float = 0.05):
    """The reference's ResultsWrong protocol (tests/testsuite.cpp:492-618):
    `run_short(i)` returns the state of the i-th short run (300 desorptions
    there); each is compared with the stored state at the tighter
    thresholds 0.006 / 0.05 and counts as detected when Compare reports a
    global or facet difference. Returns (detected, n_runs, comparable):
    the reference requires detected / n_runs >= 0.66, and treats a run
    whose hits equal its desorptions as not comparable (all detected)."""
    detected = 0
    for i in range(n_runs):
        s = run_short(i)
        if int(s.glob[0]["nb_mc_hit"]) == int(s.glob[0]["nb_desorbed"]):
            return n_runs, n_runs, False
        g, f, _, _ = compare_states(stored, s, glob=glob, loc=loc)
        if g or f:
            detected += 1
    return detected, n_runs, True


def binomial_z(k1: float, n1: float, k2: float, n2: float) -> float:
    """z-score of two binomial proportions (pooled variance)."""
    p1, p2 = k1 / n1, k2 / n2
    p = (k1 + k2) / (n1 + n2)
    s = math.sqrt(max(p * (1 - p) * (1 / n1 + 1 / n2), 1e-300))
    return (p1 - p2) / s


def proportion_z(k: float, n: float, p_ref: float, sigma_ref: float = 0.0) -> float:
    """z of a measured proportion k/n against a reference value with its own sigma."""
    p = k / n
    s = math.sqrt(max(p_ref * (1 - p_ref) / n + sigma_ref ** 2, 1e-300))
    return (p - p_ref) / s


def absorption_chi2(a1: np.ndarray, n1: float, a2: np.ndarray, n2: float, min_count: float = 20.0):
    """Two-sample chi-square over multinomial absorption counts.
    Returns (chi2, dof, p_value)."""
    from scipy import stats
    a1 = np.asarray(a1, dtype=np.float64)
    a2 = np.asarray(a2, dtype=np.float64)
    keep = (a1 + a2) >= min_count
    x, y = a1[keep], a2[keep]
    rest1, rest2 = n1 - x.sum(), n2 - y.sum()
    if rest1 + rest2 >= min_count:
        x = np.append(x, rest1); y = np.append(y, rest2)
    k1 = math.sqrt(n2 / n1)
    k2 = math.sqrt(n1 / n2)
    chi2 = float(np.sum((k1 * x - k2 * y) ** 2 / (x + y)))
    dof = len(x) - 1
    return chi2, dof, float(stats.chi2.sf(chi2, dof))


def batch_means(values: np.ndarray):
    """Mean and standard error of per-batch estimates (rows = batches)."""
    v = np.asarray(values, dtype=np.float64)
    n = v.shape[0]
    return v.mean(0), v.std(0, ddof=1) / math.sqrt(n)
