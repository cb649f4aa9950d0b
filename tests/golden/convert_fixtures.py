"""Convert the reference's test fixtures (zipped MolFlow XML files holding a
model AND its stored simulation results) into compact gzipped JSON files
used by the parity tests.

Reads data only: the XML is parsed with the standard library, following the
attribute meanings of the reference loader (src/IO/LoaderXML.cpp:31-331
geometry and settings, :366-829 results, :831-1126 facets). Run in the
build container:

    python tests/golden/convert_fixtures.py /root/reference/tests/TestCases

Each output keeps the geometry, facet parameters, settings, time-dependent
parameters (catalog parameters appended after the file's own, as the
reference's LoadParameterCatalog does), user moments and the stored results:
global counters, per-facet counters for every moment, and for a bounded
subset of facets the profiles, textures, angle maps and histograms.
"""
from __future__ import annotations

import csv
import gzip
import json
import os
import sys
import xml.etree.ElementTree as ET
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_TEX_FACETS = 8
MAX_TEX_CELLS = 4096
MAX_PROFILE_FACETS = 24


def _f(node, attr, default=0.0):
    if node is None or attr not in node.attrib:
        return default
    return float(node.attrib[attr])


def _i(node, attr, default=0):
    if node is None or attr not in node.attrib:
        return default
    v = node.attrib[attr]
    if v in ("true", "false"):
        return int(v == "true")
    return int(float(v))


def _b(node, attr, default=0):
    return _i(node, attr, default)


def load_catalog(ref_root: str):
    out = []
    cat = os.path.join(ref_root, "copy_to_build", "parameter_catalog")
    if not os.path.isdir(cat):
        return out
    for fn in sorted(os.listdir(cat)):
        if not fn.endswith(".csv"):
            continue
        xs, ys = [], []
        with open(os.path.join(cat, fn), encoding="utf-8-sig") as fh:
            for row in csv.reader(fh):
                if len(row) == 2:
                    xs.append(float(row[0])); ys.append(float(row[1]))
        pts = sorted(zip(xs, ys))
        out.append({"name": "[catalog] " + fn, "x": [p[0] for p in pts], "y": [p[1] for p in pts],
                    "log_x": False, "log_y": False})
    return out


def convert(xml_bytes: bytes, name: str, catalog) -> dict:
    root = ET.fromstring(xml_bytes)
    geo = root.find("Geometry")
    verts = []
    for v in geo.find("Vertices").findall("Vertex"):
        verts.append([float(v.attrib["x"]), float(v.attrib["y"]), float(v.attrib["z"])])
    ss = root.find("MolflowSimuSettings")
    params = []
    for p in ss.find("Parameters").findall("Parameter"):
        pts = [(float(m.attrib["t"]), float(m.attrib["value"])) for m in p.findall("Moment")]
        params.append({"name": p.attrib["name"], "x": [a for a, _ in pts], "y": [b for _, b in pts],
                       "log_x": p.attrib.get("logXinterp") == "true", "log_y": p.attrib.get("logYinterp") == "true"})
    params += catalog
    names = [p["name"] for p in params]

    def param_ref(node, name_attr):
        if node is None:
            return -1
        if name_attr in node.attrib:
            nm = node.attrib[name_attr]
            return names.index(nm) if nm else -1
        pid = _i(node, "parameterId", -1)
        return pid

    facets = {}
    polys = []
    angle_in = {}
    fnodes = geo.find("Facets").findall("Facet")
    cols = {}

    def put(k, i, val):
        cols.setdefault(k, [None] * len(fnodes))[i] = val

    for i, fn in enumerate(fnodes):
        polys.append([int(x.attrib["vertex"]) for x in fn.find("Indices").findall("Indice")])
        st, op, og = fn.find("Sticking"), fn.find("Opacity"), fn.find("Outgassing")
        te, rf, stc = fn.find("Temperature"), fn.find("Reflection"), fn.find("Structure")
        rec = fn.find("Recordings")
        tx = rec.find("Texture")
        put("sticking", i, _f(st, "constValue"))
        put("sticking_param", i, param_ref(st, "stickingParam"))
        put("opacity", i, _f(op, "constValue", 1.0))
        put("opacity_param", i, param_ref(op, "opacityParam"))
        put("is_2sided", i, _i(op, "is2sided"))
        put("temperature", i, _f(te, "value", 293.15))
        put("temperature_param", i, param_ref(te, "tempParam") if te is not None and "tempParam" in te.attrib else -1)
        put("accomodation", i, _f(te, "accFactor", 1.0))
        put("desorb_type", i, _i(og, "desType"))
        put("desorb_exponent", i, _f(og, "desExponent"))
        put("outgassing", i, _f(og, "constValue"))
        put("outgassing_param", i, param_ref(og, "OutgassingParam"))
        put("use_outgassing_map", i, _b(og, "useOutgassingFile"))
        if "diffusePart" in rf.attrib:
            put("diffuse_part", i, _f(rf, "diffusePart", 1.0))
            put("specular_part", i, _f(rf, "specularPart"))
            put("cosine_exponent", i, _f(rf, "cosineExponent"))
        else:
            t = _i(rf, "type")
            put("diffuse_part", i, 1.0 if t == 0 else 0.0)
            put("specular_part", i, 1.0 if t == 1 else 0.0)
            put("cosine_exponent", i, 0.0)
        put("enable_sojourn", i, _b(rf, "enableSojournTime"))
        put("sojourn_freq", i, _f(rf, "sojournFreq", 1e13))
        put("sojourn_e", i, _f(rf, "sojournE", 100.0))
        put("super_idx", i, _i(stc, "inStructure"))
        put("super_dest", i, _i(stc, "linksTo"))
        put("teleport_dest", i, _i(fn.find("Teleport"), "target"))
        put("is_moving", i, _b(fn.find("Motion"), "isMoving"))
        put("profile_type", i, _i(rec.find("Profile"), "type"))
        put("is_textured", i, int(tx.attrib.get("hasMesh") == "true"))
        put("tex_width_precise", i, _f(tx, "texDimX"))
        put("tex_height_precise", i, _f(tx, "texDimY"))
        for k, a in (("count_des", "countDes"), ("count_abs", "countAbs"), ("count_refl", "countRefl"),
                     ("count_trans", "countTrans"), ("count_dir", "countDir")):
            put(k, i, _b(tx, a))
        am = rec.find("IncidentAngleMap")
        put("am_record", i, _b(am, "record"))
        put("am_phi_width", i, _i(am, "phiWidth"))
        put("am_theta_limit", i, _f(am, "thetaLimit", 1.570796326))
        put("am_theta_lower_res", i, _i(am, "thetaLowerRes"))
        put("am_theta_higher_res", i, _i(am, "thetaHigherRes"))
        h = fn.find("Histograms")
        hb = h.find("Bounces") if h is not None else None
        hd = h.find("Distance") if h is not None else None
        ht = h.find("Time") if h is not None else None
        put("fh_record_bounce", i, int(hb is not None))
        put("fh_bounce_binsize", i, _i(hb, "binSize"))
        put("fh_bounce_max", i, _i(hb, "max"))
        put("fh_record_distance", i, int(hd is not None))
        put("fh_distance_binsize", i, _f(hd, "binSize"))
        put("fh_distance_max", i, _f(hd, "max"))
        put("fh_record_time", i, int(ht is not None))
        put("fh_time_binsize", i, _f(ht, "binSize"))
        put("fh_time_max", i, _f(ht, "max"))
        ia = fn.find("IncidentAngleMap")
        if ia is not None and ia.find("map") is not None and "angleMapThetaLimit" in ia.attrib:
            angle_in[i] = {"phi_width": _i(ia, "angleMapPhiWidth"), "theta_limit": _f(ia, "angleMapThetaLimit"),
                           "lower": _i(ia, "angleMapThetaLowerRes"), "higher": _i(ia, "angleMapThetaHigherRes"),
                           "map": [int(float(x)) for x in ia.find("map").text.split()]}
        elif cols["desorb_type"][i] == 4:
            cols["desorb_type"][i] = 0  # reference drops DES_ANGLEMAP without a map (LoaderXML.cpp:1071-1073)
    gas = ss.find("Gas")
    ts = ss.find("TimeSettings")
    lf = ss.find("LowFluxMode")
    sc = gas.find("Scattering") if gas is not None else None
    mo = ss.find("Motion")
    mf = ss.find("MeasureForces")
    settings = {
        "gas_mass": _f(gas, "mass", 28.0), "half_life": _f(gas, "halfLife", 1.0),
        "enable_decay": _b(gas, "enableDecay") if "enableDecay" in gas.attrib else int(_f(gas, "halfLife", 1e100) < 1e100),
        "time_window": _f(ts, "timeWindow"), "use_maxwell": _b(ts, "useMaxwellDistr"),
        "calc_constant_flow": _b(ts, "calcConstFlow"),
        "low_flux_mode": _b(lf, "enabled"), "low_flux_cutoff": _f(lf, "cutoff", 1e-7),
        "scatter_enabled": _b(sc, "enabled"), "scatter_mass_ratio": _f(sc, "massRatio"),
        "scatter_mean_free_path_cm": _f(sc, "meanFreePath_cm"),
        "motion_type": _i(mo, "type"),
        "enable_force": _b(mf, "enabled"),
    }
    if settings["motion_type"] == 1:
        v = mo.find("VelocityVector")
        settings["motion_vector2"] = [_f(v, "vx"), _f(v, "vy"), _f(v, "vz")]
    elif settings["motion_type"] == 2:
        v, v2 = mo.find("AxisBasePoint"), mo.find("RotationVector")
        settings["motion_vector1"] = [_f(v, "x"), _f(v, "y"), _f(v, "z")]
        settings["motion_vector2"] = [_f(v2, "x"), _f(v2, "y"), _f(v2, "z")]
    gh = ss.find("Global_histograms")
    if gh is not None:
        b, d, t = gh.find("Bounces"), gh.find("Distance"), gh.find("Time")
        if b is not None:
            settings.update(gh_record_bounce=1, gh_bounce_binsize=_i(b, "binSize"), gh_bounce_max=_i(b, "max"))
        if d is not None:
            settings.update(gh_record_distance=1, gh_distance_binsize=_f(d, "binSize"), gh_distance_max=_f(d, "max"))
        if t is not None:
            settings.update(gh_record_time=1, gh_time_binsize=_f(t, "binSize"), gh_time_max=_f(t, "max"))
    user_moments = []
    um = ts.find("UserMoments")
    if um is not None:
        for e in um.findall("UserEntry"):
            w = float(e.attrib.get("window", 0))
            user_moments.append([e.attrib["content"], w if w != 0 else settings["time_window"]])

    # ---- stored results
    res = root.find("MolflowResults")
    moments = res.find("Moments").findall("Moment")
    gl = moments[0].find("Global").find("Hits").attrib
    glob = {k: float(v) for k, v in gl.items()}
    fields = ("nbHit", "nbHitEquiv", "nbDes", "nbAbsEquiv", "sum_v_ort", "sum_1_per_v", "sum_v")
    nF = len(fnodes)
    per_slot = {k: [] for k in fields}
    profiles, textures, hists = {}, {}, {}
    prof_facets = [i for i in range(nF) if cols["profile_type"][i] != 0][:MAX_PROFILE_FACETS]
    tex_facets = [i for i in range(nF) if cols["is_textured"][i]][:MAX_TEX_FACETS]
    for m_i, mom in enumerate(moments):
        rows = {k: [0.0] * nF for k in fields}
        fr = mom.find("FacetResults")
        for fres in fr.findall("Facet"):
            fid = int(fres.attrib["id"])
            h = fres.find("Hits")
            for k in fields:
                rows[k][fid] = float(h.attrib.get(k, 0.0))
            if fid in prof_facets and fres.find("Profile") is not None:
                sl = fres.find("Profile").findall("Slice")
                profiles.setdefault(str(fid), []).append(
                    [[float(s.attrib.get("countEquiv", s.attrib.get("count", 0))), float(s.attrib.get("sum_v_ort", 0)),
                      float(s.attrib.get("sum_1_per_v", 0))] for s in sl])
            t = fres.find("Texture")
            if (fid in tex_facets and m_i == 0 and t is not None
                    and int(t.attrib["width"]) * int(t.attrib["height"]) <= MAX_TEX_CELLS):
                cnode = t.find("countEquiv") if t.find("countEquiv") is not None else t.find("count")
                textures[str(fid)] = {"w": int(t.attrib["width"]), "h": int(t.attrib["height"]),
                                      "count": [float(x) for x in cnode.text.split()],
                                      "sum_v_ort": [float(x) for x in t.find("sum_v_ort").text.split()],
                                      "sum_1_per_v": [float(x) for x in t.find("sum_1_per_v").text.split()]}
            hn = fres.find("Histograms")
            if hn is not None and m_i == 0:
                hh = {}
                for kind in ("Bounces", "Distance", "Time"):
                    k = hn.find(kind)
                    if k is not None:
                        hh[kind] = [float(b.attrib["count"]) for b in k.findall("Bin")]
                if hh:
                    hists[str(fid)] = hh
        for k in fields:
            per_slot[k].append(rows[k])
    ghist = {}
    hn = moments[0].find("Histograms")
    if hn is not None:
        for kind in ("Bounces", "Distance", "Time"):
            k = hn.find(kind)
            if k is not None:
                ghist[kind] = [float(b.attrib["count"]) for b in k.findall("Bin")]
    return {
        "name": name, "n_structures": int(geo.find("Structures").attrib["nb"]),
        "vertices": verts, "polygons": polys, "facets": cols, "settings": settings, "params": params,
        "user_moments": user_moments, "angle_maps_in": angle_in,
        "results": {"n_slots": len(moments), "global": glob, "facets": per_slot, "profiles": profiles,
                    "textures": textures, "facet_hist": hists, "global_hist": ghist},
    }


def main(src: str):
    ref_root = os.path.abspath(os.path.join(src, "..", ".."))
    catalog = load_catalog(ref_root)
    for fn in sorted(os.listdir(src)):
        if not fn.endswith(".zip"):
            continue
        with zipfile.ZipFile(os.path.join(src, fn)) as z:
            xmls = [n for n in z.namelist() if n.endswith(".xml")]
            data = z.read(xmls[0])
        name = fn[:-4]
        d = convert(data, name, catalog)
        d["source"] = "tests/TestCases/" + fn
        out = os.path.join(HERE, name.replace(" ", "_") + ".json.gz")
        with gzip.open(out, "wt", encoding="utf-8") as fh:
            json.dump(d, fh, separators=(",", ":"))
        print(f"{name}: {len(d['polygons'])} facets, {d['results']['n_slots']} slots -> {os.path.getsize(out)} B")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/TestCases")
