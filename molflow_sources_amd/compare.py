"""Result comparison: the reference's parity checker and sigma-based tests.

compare_states restates GlobalSimuState::Compare
(reference src/Simulation/MolflowSimGeom.cpp:558-1148) at the global and
facet levels: relative thresholds `glob` (0.009 in the reference test) and
`loc` (0.07), per-facet skip below 40000 MC hits, adaptive local threshold
max(loc, 1/sqrt(nMC)), value checks only when sqrt(min count) >= 80, and a
40x looser threshold on the 1/velocity sums. IsEqual(a, b, tol), absent
from the checkout, is taken as |a-b| <= tol * max(|a|,|b|).

The sigma-based helpers add what the reference lacks: a binomial z-score
for transmission-type ratios, a multinomial chi-square over per-facet
absorption counts (each particle is absorbed once), and batch-means
standard errors for correlated hit tallies.
"""
from __future__ import annotations

import math

import numpy as np


def is_equal(a: float, b: float, tol: float) -> bool:
    return abs(a - b) <= tol * max(abs(a), abs(b)) or (a == b)


def compare_states(lhs, rhs, glob: float = 0.009, loc: float = 0.07, moments: bool = True):
    """Return (n_global_errors, n_facet_errors, messages). lhs/rhs expose
    .glob (global counters) and .facets[slot][facet] counters."""
    msgs = []
    g_err = f_err = 0
    L, R = lhs.glob[0], rhs.glob[0]
    if L["nb_desorbed"] == 0 or R["nb_desorbed"] == 0:
        return 1, -1, ["no desorptions"]
    a1, a2 = L["nb_abs_equiv"] / L["nb_desorbed"], R["nb_abs_equiv"] / R["nb_desorbed"]
    if not is_equal(a1, a2, glob):
        g_err += 1; msgs.append(f"[Global][abs/des] {a1} vs {a2}")
    h1, h2 = L["nb_mc_hit"] / L["nb_desorbed"], R["nb_mc_hit"] / R["nb_desorbed"]
    if not is_equal(h1, h2, glob):
        g_err += 1; msgs.append(f"[Global][hits/des] {h1} vs {h2}")
    slots = range(lhs.facets.shape[0] if moments else 1)
    for m in slots:
        fl, fr = lhs.facets[m], rhs.facets[m]
        for f in range(min(len(fl), len(fr))):
            cl, cr = fl[f], fr[f]
            if cl["nb_mc_hit"] < 40000 or cr["nb_mc_hit"] < 40000:
                continue
            nmc = min(cl["nb_mc_hit"], cr["nb_mc_hit"])
            lt = max(loc, 1.0 / math.sqrt(nmc))
            sc, scr = 1.0 / L["nb_hit_equiv"], 1.0 / R["nb_hit_equiv"]
            fs = 1.0 / (L["nb_hit_equiv"] + L["nb_abs_equiv"] + L["nb_desorbed"])
            fsr = 1.0 / (R["nb_hit_equiv"] + R["nb_abs_equiv"] + R["nb_desorbed"])
            if math.sqrt(max(1.0, min(cl["nb_hit_equiv"], cr["nb_hit_equiv"]))) >= 80:
                pairs = [("hits", cl["nb_hit_equiv"] * sc, cr["nb_hit_equiv"] * scr, lt),
                         ("sum_v_ort", cl["sum_v_ort"] * sc, cr["sum_v_ort"] * scr, lt),
                         ("sum_1_per_velocity", cl["sum_1_per_velocity"] * fs, cr["sum_1_per_velocity"] * fsr, lt * 40),
                         ("sum_1_per_ort_velocity", cl["sum_1_per_ort_velocity"] * fs,
                          cr["sum_1_per_ort_velocity"] * fsr, lt * 40)]
                for name, x, y, tol in pairs:
                    if not is_equal(x, y, tol):
                        f_err += 1; msgs.append(f"[Facet][{f}][{name}][moment {m}] {x} vs {y}")
            if math.sqrt(max(1.0, min(cl["nb_abs_equiv"], cr["nb_abs_equiv"]))) >= 80:
                x, y = cl["nb_abs_equiv"] / cl["nb_mc_hit"], cr["nb_abs_equiv"] / cr["nb_mc_hit"]
                if not is_equal(x, y, lt):
                    f_err += 1; msgs.append(f"[Facet][{f}][abs/hits][moment {m}] {x} vs {y}")
            if math.sqrt(max(1.0, min(cl["nb_desorbed"], cr["nb_desorbed"]))) >= 80:
                x, y = cl["nb_desorbed"] / L["nb_desorbed"], cr["nb_desorbed"] / R["nb_desorbed"]
                if not is_equal(x, y, lt):
                    f_err += 1; msgs.append(f"[Facet][{f}][des][moment {m}] {x} vs {y}")
    return g_err, f_err, msgs


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
