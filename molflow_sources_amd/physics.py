"""Physical quantities from the tallies: the MoleculesPerTP scaling, facet
pressure, density, impingement rate and mean speed, the CSV facet exports
of molflowCLI, and the formula variables of MolFlow+'s formula editor.

Formulas follow the reference:
  GetMoleculesPerTP       src/IO/CSVExporter.cpp:59-79
  DensityCorrection       src/IO/CSVExporter.cpp:84-101
  GetArea                 src/IO/CSVExporter.cpp:104-106 (x2 when 2-sided)
  facet quantities        src/IO/CSVExporter.cpp:273-323
  facet CSV exports       src/IO/CSVExporter.cpp:341-430
  formula variables       src/FormulaEvaluator_MF.cpp:25-250

The model totals (final outgassing rate, total desorbed molecules, latest
moment) are the ones the library derives when it prepares a model
(CalcTotalOutgassing, MolflowSimGeom.cpp:290-335), read through
mfg_model_totals, which needs no device.
"""
from __future__ import annotations

import ast
import ctypes as C
import math
import operator
from dataclasses import dataclass

import numpy as np

AVOGADRO_REF = 6e23      # the reference's Avogadro constant in these formulas
KB_FORMULA = 1.3806504e-23


@dataclass
class ModelTotals:
    final_outgassing_rate: float           # molecules/s (sp.finalOutgassingRate)
    final_outgassing_rate_pa_m3_s: float   # Pa m3/s (sp.finalOutgassingRate_Pa_m3_sec)
    total_desorbed_molecules: float        # sp.totalDesorbedMolecules
    latest_moment: float                   # s


def model_totals(model) -> ModelTotals:
    from .simulation import MolflowGpuError, lib
    v = model.view()
    out = (C.c_double * 4)()
    rc = lib().mfg_model_totals(C.byref(v), out)
    if rc:
        raise MolflowGpuError(f"mfg_model_totals failed ({rc}): the model does not load")
    return ModelTotals(out[0], out[1], out[2], out[3])


def sorted_moments(model):
    """(time, window) of the moments in slot order (slot m = moment m-1)."""
    return sorted(model.moments, key=lambda mw: mw[0] - 0.5 * mw[1])


def molecules_per_tp(moment: int, totals: ModelTotals, n_desorbed: int, model) -> float:
    """Real molecules one test particle stands for (GetMoleculesPerTP):
    steady state final outgassing rate / desorbed; moment m total desorbed
    molecules / window_m / desorbed."""
    if n_desorbed == 0:
        return 0.0
    if moment == 0:
        return totals.final_outgassing_rate / float(n_desorbed)
    window = sorted_moments(model)[moment - 1][1]
    return totals.total_desorbed_molecules / window / float(n_desorbed)


def facet_areas(model) -> np.ndarray:
    """GetArea: the polygon area (cm^2), doubled for 2-sided facets."""
    return model.facet_areas() * np.where(model.facets["is_2sided"] != 0, 2.0, 1.0)


def density_correction(counters) -> np.ndarray:
    """Correction of the double-density effect on desorbing / absorbing
    facets (DensityCorrection), per facet."""
    c = counters
    hit_or_des = (c["nb_mc_hit"] > 0) | (c["nb_desorbed"] > 0)
    abs_or_des = (c["nb_abs_equiv"] > 0) | (c["nb_desorbed"] > 0)
    den = c["nb_hit_equiv"] + c["nb_desorbed"].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        corr = 1.0 - (c["nb_abs_equiv"] + c["nb_desorbed"].astype(np.float64)) / den / 2.0
    return np.where(hit_or_des & abs_or_des, corr, 1.0)


def facet_quantities(model, tallies, totals: ModelTotals, moment: int = 0) -> dict:
    """Per-facet physical quantities of one moment slot (CSVExporter
    F_IMPINGEMENT, F_DENSITY1P, F_DENSITYKGP, F_PRESSURE, F_AVGSPEED):
      impingement  hits / s / m^2
      density      1 / m^3
      density_kg   kg / m^3
      pressure     mbar
      avg_speed    m / s"""
    c = tallies.facets[moment]
    n_des = int(tallies.glob[0]["nb_desorbed"])
    mptp = molecules_per_tp(moment, totals, n_des, model)
    area = facet_areas(model)
    mass = model.settings.gas_mass
    dc = density_correction(c)
    with np.errstate(divide="ignore", invalid="ignore"):
        imp = c["nb_hit_equiv"] / area * 1e4 * mptp
        dens = c["sum_1_per_ort_velocity"] / area * 1e4 * mptp * dc
        dens_kg = dens * mass / 1000.0 / AVOGADRO_REF
        pres = c["sum_v_ort"] * 1e4 * mptp * (mass / 1000.0 / AVOGADRO_REF) * 0.01 / area
        spd = np.where(c["sum_1_per_velocity"] == 0.0, 0.0,
                       (c["nb_hit_equiv"] + c["nb_desorbed"].astype(np.float64)) / c["sum_1_per_velocity"])
    return {"impingement": imp, "density": dens, "density_kg": dens_kg, "pressure": pres, "avg_speed": spd}


def write_facet_quantities_csv(path: str, model, tallies, totals: ModelTotals) -> None:
    """facet_physics.csv of molflowCLI --writeFacetQuantities
    (ExportPhysicalQuantitiesForFacets): steady-state slot."""
    q = facet_quantities(model, tallies, totals, 0)
    cols = [("#", None), ("Imping.rate", "impingement"), ("Pressure [mbar]", "pressure"),
            ("Density [1/m3]", "density"), ("Density [kg/m3]", "density_kg"), ("Av.mol.speed[m/s]", "avg_speed")]
    with open(path, "w") as fh:
        fh.write(",".join(h for h, _ in cols) + "\n")
        for f in range(model.n_facets):
            fh.write(",".join(str(f + 1) if k is None else repr(float(q[k][f])) for _, k in cols) + "\n")


DES_NAMES = ("None", "Uniform", "Cosine", "Cosine^", "Angle map")
PROFILE_NAMES = ("None", "Pressure (u)", "Pressure (v)", "Angular", "Speed distr.", "Ort. velocity", "Tan. velocity")


def write_facet_details_csv(path: str, model, tallies, totals: ModelTotals) -> None:
    """facet_details.csv of molflowCLI --writeFacetDetails
    (ExportAllFacetDetails): facet settings, counters and quantities."""
    from .model import texture_cells
    F = model.facets
    c = tallies.facets[0]
    q = facet_quantities(model, tallies, totals, 0)
    area = model.facet_areas()
    hdr = ["#", "Sticking", "Opacity", "Structure", "Link", "Desorption", "Reflection", "2 Sided", "Vertex", "Area",
           "Temperature (K)", "Texture (u x v)", "Profile", "Imping.rate", "Density [1/m3]", "Density [kg/m3]",
           "Pressure [mbar]", "Av.mol.speed[m/s]", "MC Hits", "Equiv.hits", "Des.", "Equiv.abs."]
    with open(path, "w") as fh:
        fh.write(",".join(hdr) + "\n")
        for f in range(model.n_facets):
            dt = int(F["desorb_type"][f])
            des = DES_NAMES[dt] if dt < len(DES_NAMES) else str(dt)
            if dt == 3:
                des += repr(float(F["desorb_exponent"][f]))
            refl = (f"diff. {F['diffuse_part'][f]:.3g}, spec. {F['specular_part'][f]:.3g}, "
                    f"cos^{F['cosine_exponent'][f]:.3g}").replace(",", ";")
            tex = (f"{texture_cells(F['tex_width_precise'][f])}x{texture_cells(F['tex_height_precise'][f])}"
                   if F["is_textured"][f] else "None")
            row = [str(f + 1), repr(float(F["sticking"][f])), repr(float(F["opacity"][f])),
                   str(int(F["super_idx"][f]) + 1), str(int(F["super_dest"][f])), des, refl,
                   "Yes" if F["is_2sided"][f] else "No", str(len(model.polygons[f])),
                   (f"2*{area[f]!r}" if F["is_2sided"][f] else repr(float(area[f]))), repr(float(F["temperature"][f])),
                   tex, PROFILE_NAMES[int(F["profile_type"][f])],
                   repr(float(q["impingement"][f])), repr(float(q["density"][f])), repr(float(q["density_kg"][f])),
                   repr(float(q["pressure"][f])), repr(float(q["avg_speed"][f])), str(int(c["nb_mc_hit"][f])),
                   repr(float(c["nb_hit_equiv"][f])), str(int(c["nb_desorbed"][f])), repr(float(c["nb_abs_equiv"][f]))]
            fh.write(",".join(row) + "\n")


# ---------------------------------------------------------------- formulas
_FACET_VARS = ("MCH", "DEN", "AR", "A", "D", "H", "P", "Z", "V", "T")


def formula_variables(model, tallies, totals: ModelTotals, moment: int = 0) -> dict:
    """The variables of MolFlow+'s formula editor for one moment slot:
    An, Dn, MCHn, Hn, Pn, DENn, Zn, Vn, Tn, ARn (facet n, 1-based) and the
    global SUMDES, SUMABS, SUMMCHIT, SUMHIT, MPP, MFP, DESAR, ABSAR, QCONST,
    QCONST_N, NTOT, GASMASS, KB, R, Na, PI (FormulaEvaluator_MF.cpp)."""
    c = tallies.facets[moment]
    g = tallies.glob[0]
    q = facet_quantities(model, tallies, totals, moment)
    area_get = facet_areas(model)
    area_sh = model.facet_areas()
    F = model.facets
    v = {}
    for f in range(model.n_facets):
        n = f + 1
        v[f"A{n}"] = float(c["nb_abs_equiv"][f])
        v[f"D{n}"] = float(c["nb_desorbed"][f])
        v[f"MCH{n}"] = float(c["nb_mc_hit"][f])
        v[f"H{n}"] = float(c["nb_hit_equiv"][f])
        v[f"P{n}"] = float(q["pressure"][f])
        v[f"DEN{n}"] = float(q["density"][f])
        v[f"Z{n}"] = float(q["impingement"][f])
        v[f"V{n}"] = float(q["avg_speed"][f])
        v[f"T{n}"] = float(F["temperature"][f])
        v[f"AR{n}"] = float(area_sh[f])
    v["SUMDES"] = float(g["nb_desorbed"])
    v["SUMABS"] = float(g["nb_abs_equiv"])
    v["SUMMCHIT"] = float(g["nb_mc_hit"])
    v["SUMHIT"] = float(g["nb_hit_equiv"])
    v["MPP"] = float(g["dist_traveled_total"]) / float(g["nb_desorbed"]) if g["nb_desorbed"] else math.nan
    v["MFP"] = float(g["dist_traveled_full_hits_only"]) / float(g["nb_hit_equiv"]) if g["nb_hit_equiv"] else math.nan
    v["DESAR"] = float(area_get[F["desorb_type"] != 0].sum())
    v["ABSAR"] = float((area_get * F["opacity"])[F["sticking"] > 0].sum())
    v["QCONST"] = totals.final_outgassing_rate_pa_m3_s * 10.0
    v["QCONST_N"] = totals.final_outgassing_rate
    v["NTOT"] = totals.total_desorbed_molecules
    v["GASMASS"] = float(model.settings.gas_mass)
    v["KB"] = KB_FORMULA
    v["R"] = 8.314472
    v["Na"] = 6.02214179e23
    v["PI"] = math.pi
    return v


_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}
_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "log10": math.log10, "sin": math.sin, "cos": math.cos,
          "tan": math.tan, "abs": abs, "asin": math.asin, "acos": math.acos, "atan": math.atan}


def evaluate_formula(expr: str, variables: dict) -> float:
    """Evaluate an arithmetic formula over the variables (e.g. "A2/SUMDES",
    the transmission of a pipe desorbing from facet 1 into facet 2).
    Only numbers, variables, + - * / ^ and a few math functions are
    accepted; variable names are case-insensitive."""
    low = {k.lower(): val for k, val in variables.items()}
    tree = ast.parse(expr.replace("^", "**"), mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name):
            if n.id.lower() not in low:
                raise KeyError(f"unknown variable {n.id}")
            return low[n.id.lower()]
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.operand))
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id.lower() in _FUNCS and len(n.args) == 1:
            return _FUNCS[n.func.id.lower()](ev(n.args[0]))
        raise ValueError(f"unsupported expression: {ast.dump(n)}")

    return float(ev(tree))
