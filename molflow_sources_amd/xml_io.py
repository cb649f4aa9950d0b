"""MolFlow+ XML model and result files: the XmlLoader / XmlWriter roles
(reference src/IO/LoaderXML.cpp, src/IO/WriterXML.cpp) for the GPU unit.

    doc = xml_io.load("model.xml")        # or a .zip holding one .xml
    sim.load(doc.model)                   # molflow_sources_amd.model.Model
    ...
    xml_io.save("result.xml", doc, tallies, des_limit=n)

load() reads the geometry, the facet parameters (LoaderXML.cpp:831-1126),
the simulation settings and time-dependent parameters (:31-331), and, when
the file holds them, the stored results (:366-829) as a Tallies object laid
out like the library's tally buffers (include/molflow_gpu.h), so a run can
resume from them. save() writes the document back with its MolflowResults
replaced by the given tallies, in the layout of
XmlWriter::SaveSimulationState (WriterXML.cpp:367-677), and the facets'
recorded incident-angle maps updated; MolFlow+ opens such a file with its
results. Only the standard library's XML parser is used.
"""
from __future__ import annotations

import csv
import ctypes
import io
import math
import os
import xml.etree.ElementTree as ET
import zipfile
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .model import Model, Parameter, default_facets, default_settings, texture_cells

PROFILE_SIZE = A.PROFILE_SIZE


class XmlFormatError(ValueError):
    pass


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


def read_document(path_or_bytes) -> ET.Element:
    """Root element of a MolFlow XML file, a .zip holding one, or bytes."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as fh:
            data = fh.read()
    if data[:2] == b"PK":
        with zipfile.ZipFile(io.BytesIO(data)) as z:
            names = [n for n in z.namelist() if n.lower().endswith(".xml")]
            if not names:
                raise XmlFormatError("zip archive holds no .xml file")
            data = z.read(names[0])
    root = ET.fromstring(data)
    if root.find("Geometry") is None:
        raise XmlFormatError("not a MolFlow XML file (no Geometry node)")
    return root


def load_parameter_catalog(directory: str | None) -> list[Parameter]:
    """The parameter catalog (LoadParameterCatalog, MolflowSimGeom.cpp:1250-1320):
    every two-column CSV of the directory becomes a parameter named
    "[catalog] <file name>", appended after the file's own parameters."""
    out: list[Parameter] = []
    if not directory or not os.path.isdir(directory):
        return out
    for fn in sorted(os.listdir(directory)):
        if not fn.endswith(".csv"):
            continue
        pts = []
        with open(os.path.join(directory, fn), encoding="utf-8-sig") as fh:
            for row in csv.reader(fh):
                if len(row) == 2:
                    pts.append((float(row[0]), float(row[1])))
        pts.sort()
        out.append(Parameter("[catalog] " + fn, [p[0] for p in pts], [p[1] for p in pts]))
    return out


@dataclass
class XmlDocument:
    """A loaded MolFlow XML file: the parsed tree (kept for writing back),
    the model and, if the file holds results, those results."""
    root: ET.Element
    model: Model
    results: object | None = None            # Tallies, or None
    des_limit: int = 0                       # maxDesorption stored with the results
    user_moments: list = field(default_factory=list)


# ---------------------------------------------------------------- facets
def _facet_fields(fn: ET.Element, param_ref) -> dict:
    st, op, og = fn.find("Sticking"), fn.find("Opacity"), fn.find("Outgassing")
    te, rf, stc = fn.find("Temperature"), fn.find("Reflection"), fn.find("Structure")
    rec = fn.find("Recordings")
    tx = rec.find("Texture") if rec is not None else None
    d = {
        "sticking": _f(st, "constValue"), "sticking_param": param_ref(st, "stickingParam"),
        "opacity": _f(op, "constValue", 1.0), "opacity_param": param_ref(op, "opacityParam"),
        "is_2sided": _i(op, "is2sided"),
        "temperature": _f(te, "value", 293.15),
        "temperature_param": param_ref(te, "tempParam") if te is not None and "tempParam" in te.attrib else -1,
        "accomodation": _f(te, "accFactor", 1.0),
        "desorb_type": _i(og, "desType"), "desorb_exponent": _f(og, "desExponent"),
        "outgassing": _f(og, "constValue"), "outgassing_param": param_ref(og, "OutgassingParam"),
        "use_outgassing_map": _i(og, "useOutgassingFile"),
        "enable_sojourn": _i(rf, "enableSojournTime"), "sojourn_freq": _f(rf, "sojournFreq", 1e13),
        "sojourn_e": _f(rf, "sojournE", 100.0),
        "super_idx": _i(stc, "inStructure"), "super_dest": _i(stc, "linksTo"),
        "teleport_dest": _i(fn.find("Teleport"), "target"), "is_moving": _i(fn.find("Motion"), "isMoving"),
        "profile_type": _i(rec.find("Profile") if rec is not None else None, "type"),
        "is_textured": int(tx is not None and tx.attrib.get("hasMesh") == "true"),
        "tex_width_precise": _f(tx, "texDimX"), "tex_height_precise": _f(tx, "texDimY"),
    }
    # ReflectionParam: the diffuse / specular / cos^N split (MolflowTypes.h:230-241),
    # or the older single reflection type
    if rf is not None and "diffusePart" in rf.attrib:
        d.update(diffuse_part=_f(rf, "diffusePart", 1.0), specular_part=_f(rf, "specularPart"),
                 cosine_exponent=_f(rf, "cosineExponent"))
    else:
        t = _i(rf, "type")
        d.update(diffuse_part=1.0 if t == 0 else 0.0, specular_part=1.0 if t == 1 else 0.0, cosine_exponent=0.0)
    for k, a in (("count_des", "countDes"), ("count_abs", "countAbs"), ("count_refl", "countRefl"),
                 ("count_trans", "countTrans"), ("count_dir", "countDir")):
        d[k] = _i(tx, a)
    am = rec.find("IncidentAngleMap") if rec is not None else None
    d.update(am_record=_i(am, "record"), am_phi_width=_i(am, "phiWidth"),
             am_theta_limit=_f(am, "thetaLimit", 1.570796326), am_theta_lower_res=_i(am, "thetaLowerRes"),
             am_theta_higher_res=_i(am, "thetaHigherRes"))
    h = fn.find("Histograms")
    hb = h.find("Bounces") if h is not None else None
    hd = h.find("Distance") if h is not None else None
    ht = h.find("Time") if h is not None else None
    d.update(fh_record_bounce=int(hb is not None), fh_bounce_binsize=_i(hb, "binSize"), fh_bounce_max=_i(hb, "max"),
             fh_record_distance=int(hd is not None), fh_distance_binsize=_f(hd, "binSize"),
             fh_distance_max=_f(hd, "max"), fh_record_time=int(ht is not None), fh_time_binsize=_f(ht, "binSize"),
             fh_time_max=_f(ht, "max"))
    return d


def _settings(ss: ET.Element):
    s = default_settings()
    gas, ts = ss.find("Gas"), ss.find("TimeSettings")
    s.gas_mass = _f(gas, "mass", 28.0)
    s.half_life = _f(gas, "halfLife", 1.0)
    s.enable_decay = _i(gas, "enableDecay") if gas is not None and "enableDecay" in gas.attrib else \
        int(_f(gas, "halfLife", 1e100) < 1e100)
    sc = gas.find("Scattering") if gas is not None else None
    if sc is not None:
        s.scatter_enabled = _i(sc, "enabled")
        s.scatter_mass_ratio = _f(sc, "massRatio")
        s.scatter_mean_free_path_cm = _f(sc, "meanFreePath_cm")
        cut = sc.find("LowSpeedCutoff")
        if cut is not None:
            s.scatter_enable_cutoff = _i(cut, "enabled")
            s.scatter_cutoff_speed = _f(cut, "speed")
    lf = ss.find("LowFluxMode")
    if lf is not None:
        s.low_flux_mode = _i(lf, "enabled")
        s.low_flux_cutoff = _f(lf, "cutoff", 1e-7)
    s.time_window = _f(ts, "timeWindow")
    s.use_maxwell = _i(ts, "useMaxwellDistr")
    s.calc_constant_flow = _i(ts, "calcConstFlow")
    mo = ss.find("Motion")
    s.motion_type = _i(mo, "type")
    if s.motion_type == 1:
        v = mo.find("VelocityVector")
        for k, a in enumerate(("vx", "vy", "vz")):
            s.motion_vector2[k] = _f(v, a)
    elif s.motion_type == 2:
        v, v2 = mo.find("AxisBasePoint"), mo.find("RotationVector")
        for k, a in enumerate("xyz"):
            s.motion_vector1[k] = _f(v, a)
            s.motion_vector2[k] = _f(v2, a)
    mf = ss.find("MeasureForces")
    if mf is not None:
        s.enable_force = _i(mf, "enabled")
        tq = mf.find("Torque")
        rp = tq.find("refPoint") if tq is not None else None
        for k, a in enumerate("xyz"):
            s.torque_ref[k] = _f(rp, a)
    gh = ss.find("Global_histograms")
    if gh is not None:
        b, d, t = gh.find("Bounces"), gh.find("Distance"), gh.find("Time")
        if b is not None:
            s.gh_record_bounce, s.gh_bounce_binsize, s.gh_bounce_max = 1, _i(b, "binSize"), _i(b, "max")
        if d is not None:
            s.gh_record_distance, s.gh_distance_binsize, s.gh_distance_max = 1, _f(d, "binSize"), _f(d, "max")
        if t is not None:
            s.gh_record_time, s.gh_time_binsize, s.gh_time_max = 1, _f(t, "binSize"), _f(t, "max")
    return s


def load(path_or_bytes, catalog_dir: str | None = None, with_results: bool = True) -> XmlDocument:
    """Load a MolFlow XML (or zipped XML) file into a Model, and its stored
    results if present and asked for."""
    root = read_document(path_or_bytes)
    geo = root.find("Geometry")
    ss = root.find("MolflowSimuSettings")
    if ss is None:
        raise XmlFormatError("no MolflowSimuSettings node")
    verts = np.array([[float(v.attrib["x"]), float(v.attrib["y"]), float(v.attrib["z"])]
                      for v in geo.find("Vertices").findall("Vertex")], dtype=np.float64)
    params = []
    pnode = ss.find("Parameters")
    for p in (pnode.findall("Parameter") if pnode is not None else []):
        pts = [(float(m.attrib["t"]), float(m.attrib["value"])) for m in p.findall("Moment")]
        params.append(Parameter(p.attrib["name"], [a for a, _ in pts], [b for _, b in pts],
                                p.attrib.get("logXinterp") == "true", p.attrib.get("logYinterp") == "true"))
    params += load_parameter_catalog(catalog_dir)
    names = [p.name for p in params]

    def param_ref(node, name_attr):
        if node is None:
            return -1
        if name_attr in node.attrib:
            nm = node.attrib[name_attr]
            if not nm:
                return -1
            if nm not in names:
                raise XmlFormatError(f"unknown time-dependent parameter {nm!r}")
            return names.index(nm)
        return _i(node, "parameterId", -1)

    fnodes = geo.find("Facets").findall("Facet")
    F = default_facets(len(fnodes))
    polys = []
    og_maps, am_maps = {}, {}
    for i, fn in enumerate(fnodes):
        polys.append([int(x.attrib["vertex"]) for x in fn.find("Indices").findall("Indice")])
        for k, v in _facet_fields(fn, param_ref).items():
            F[k][i] = v
        # recorded incident-angle map: the DES_ANGLEMAP source law (the map's
        # own shape replaces the recording parameters, LoaderXML.cpp:1043-1049)
        ia = fn.find("IncidentAngleMap")
        if ia is not None and ia.find("map") is not None and "angleMapThetaLimit" in ia.attrib:
            F["am_phi_width"][i] = _i(ia, "angleMapPhiWidth")
            F["am_theta_limit"][i] = _f(ia, "angleMapThetaLimit")
            F["am_theta_lower_res"][i] = _i(ia, "angleMapThetaLowerRes")
            F["am_theta_higher_res"][i] = _i(ia, "angleMapThetaHigherRes")
            # kept for every facet that holds one (the library reads the
            # maps of DES_ANGLEMAP facets; the rest resume recordings)
            am_maps[i] = np.array([int(float(x)) for x in (ia.find("map").text or "").split()], dtype=np.uint64)
        elif F["desorb_type"][i] == A.DES_ANGLEMAP:
            F["desorb_type"][i] = A.DES_NONE        # LoaderXML.cpp:1071-1073
        # outgassing map (DynamicOutgassing, :993-1037)
        on = fn.find("DynamicOutgassing")
        if F["use_outgassing_map"][i] and on is not None and on.find("map") is not None:
            w, h = _i(on, "width"), _i(on, "height")
            ru = _f(on, "ratioU", _f(on, "ratio"))
            rv = _f(on, "ratioV", _f(on, "ratio"))
            vals = np.array([float(x) for x in (on.find("map").text or "").split()], dtype=np.float64)
            if vals.size < w * h:
                raise XmlFormatError(f"facet {i + 1}: outgassing map holds {vals.size} of {w * h} cells")
            og_maps[i] = vals[: w * h].reshape(h, w)
            F["og_map_width"][i], F["og_map_height"][i] = w, h
            F["og_ratio_u"][i], F["og_ratio_v"][i] = ru, rv
            F["total_outgassing"][i] = _f(on, "totalOutgassing")
        else:
            F["use_outgassing_map"][i] = 0
    settings = _settings(ss)
    st = geo.find("Structures")
    m = Model(verts, polys, F, settings=settings, n_structures=_i(st, "nb", 1) if st is not None else 1,
              name=root.attrib.get("name", "model"))
    m.params = params
    ts = ss.find("TimeSettings")
    user = []
    um = ts.find("UserMoments") if ts is not None else None
    if um is not None:
        for e in um.findall("UserEntry"):
            w = float(e.attrib.get("window", 0))
            user.append((e.attrib["content"], w if w != 0 else settings.time_window))
    for content, window in user:
        m.add_user_moments(content, window)
    m.og_maps, m.am_maps = og_maps, am_maps
    doc = XmlDocument(root, m, user_moments=user)
    res = root.find("MolflowResults")
    if with_results and res is not None:
        doc.results, doc.des_limit = _load_results(res, m)
        # recorded incident-angle maps live with the facets, not per moment
        _angle_maps_from_tree(root, m, doc.results)
    return doc


# ---------------------------------------------------------------- results
def tally_layout(model: Model):
    """The library's tally layout for `model` (the layout rules of
    include/molflow_gpu.h): (TallyLayout, per-facet offsets)."""
    F, s = model.facets, model.settings
    n = model.n_facets
    off = {k: np.full(n, -1, dtype=np.int64) for k in ("texture", "direction", "profile", "anglemap", "facet_hist")}
    n_tex = n_dir = n_prof = n_am = n_fh = 0
    for f in range(n):
        if F["is_textured"][f] and F["tex_width_precise"][f] > 0 and F["tex_height_precise"][f] > 0:
            cells = texture_cells(F["tex_width_precise"][f]) * texture_cells(F["tex_height_precise"][f])
            off["texture"][f] = n_tex
            if F["count_dir"][f]:
                off["direction"][f] = n_dir
                n_dir += cells
            n_tex += cells
        if F["profile_type"][f]:
            off["profile"][f] = n_prof
            n_prof += PROFILE_SIZE
        if F["am_record"][f]:
            off["anglemap"][f] = n_am
            n_am += int(F["am_phi_width"][f]) * int(F["am_theta_lower_res"][f] + F["am_theta_higher_res"][f])
        nb, nd, nt = facet_hist_sizes(F[f])
        if nb + nd + nt:
            off["facet_hist"][f] = n_fh
            n_fh += nb + nd + nt
    L = A.TallyLayout()
    L.struct_size = ctypes.sizeof(A.TallyLayout)
    L.n_facets, L.n_slots, L.n_structures = n, 1 + len(model.moments), model.n_structures
    L.n_tex_cells, L.n_dir_cells, L.n_profile_slices, L.n_angle_cells = n_tex, n_dir, n_prof, n_am
    L.n_global_hist_bins = sum(global_hist_sizes(s))
    L.n_facet_hist_bins = n_fh
    L.log_capacity = s.log_limit if s.log_enabled else 0
    L.leak_capacity = 2048
    return L, off


def _hist_sizes(rb, bb, bm, rd, db, dm, rt, tb, tm):
    nb = int(bm // bb + 2) if rb and bb else 0
    nd = int(int(dm / db) + 1) if rd and db > 0 else 0
    nt = int(int(tm / tb) + 1) if rt and tb > 0 else 0
    return nb, nd, nt


def facet_hist_sizes(f):
    return _hist_sizes(f["fh_record_bounce"], f["fh_bounce_binsize"], f["fh_bounce_max"], f["fh_record_distance"],
                       f["fh_distance_binsize"], f["fh_distance_max"], f["fh_record_time"], f["fh_time_binsize"],
                       f["fh_time_max"])


def global_hist_sizes(s):
    return _hist_sizes(s.gh_record_bounce, s.gh_bounce_binsize, s.gh_bounce_max, s.gh_record_distance,
                       s.gh_distance_binsize, s.gh_distance_max, s.gh_record_time, s.gh_time_binsize, s.gh_time_max)


def _numbers(node) -> np.ndarray:
    return np.array([float(x) for x in (node.text or "").replace(",", " ").split()], dtype=np.float64)


def _hist_into(node, dst, sizes):
    if node is None:
        return
    o = 0
    for kind, n in zip(("Bounces", "Distance", "Time"), sizes):
        k = node.find(kind)
        if k is not None and n:
            vals = [float(b.attrib["count"]) for b in k.findall("Bin")][:n]
            dst[o:o + len(vals)] = vals
        o += n


def _load_results(res: ET.Element, model: Model):
    """Stored results (LoadSimulationState, LoaderXML.cpp:366-829) as Tallies."""
    from .tallies import Tallies
    L, off = tally_layout(model)
    t = Tallies(L, off)
    moms = res.find("Moments").findall("Moment")
    if len(moms) != L.n_slots:
        raise XmlFormatError(f"results hold {len(moms)} moments, the model {L.n_slots}")
    F = model.facets
    gsz = global_hist_sizes(model.settings)
    des_limit = 0
    for m, mom in enumerate(moms):
        if m == 0:
            h = mom.find("Global").find("Hits")
            g = t.glob[0]
            g["nb_mc_hit"] = int(float(h.attrib.get("totalHit", 0)))
            g["nb_hit_equiv"] = float(h.attrib.get("totalHitEquiv", g["nb_mc_hit"]))
            g["nb_desorbed"] = int(float(h.attrib.get("totalDes", 0)))
            g["nb_abs_equiv"] = float(h.attrib.get("totalAbsEquiv", h.attrib.get("totalAbs", 0)))
            dist = float(h.attrib.get("totalDist", 0))
            g["dist_traveled_total"] = float(h.attrib.get("totalDist_total", dist))
            g["dist_traveled_full_hits_only"] = float(h.attrib.get("totalDist_fullHitsOnly", dist))
            g["nb_leak_total"] = int(float(h.attrib.get("totalLeak", 0)))
            des_limit = int(float(h.attrib.get("maxDesorption", 0)))
        _hist_into(mom.find("Histograms"), t.global_hist[m], gsz)
        fr = mom.find("FacetResults")
        for fres in (fr.findall("Facet") if fr is not None else []):
            f = int(fres.attrib["id"])
            h = fres.find("Hits")
            c = t.facets[m][f]
            c["nb_mc_hit"] = int(float(h.attrib.get("nbHit", 0)))
            c["nb_hit_equiv"] = float(h.attrib.get("nbHitEquiv", c["nb_mc_hit"]))
            c["nb_desorbed"] = int(float(h.attrib.get("nbDes", 0)))
            c["nb_abs_equiv"] = float(h.attrib.get("nbAbsEquiv", h.attrib.get("nbAbs", 0)))
            c["sum_v_ort"] = float(h.attrib.get("sum_v_ort", 0))
            c["sum_1_per_ort_velocity"] = float(h.attrib.get("sum_1_per_v", 0))
            c["sum_1_per_velocity"] = float(h.attrib.get("sum_v", 0))
            fo = fres.find("Forces")
            if fo is not None:
                for k, tag in enumerate(("Impulse", "Impulse_square", "Impulse_momentum")):
                    e = fo.find(tag)
                    for j, a in enumerate("xyz"):
                        t.forces[m][f][3 * k + j] = _f(e, a)
            pr = fres.find("Profile")
            if pr is not None and off["profile"][f] >= 0:
                p = t.profiles[m][off["profile"][f]:off["profile"][f] + PROFILE_SIZE]
                for s in pr.findall("Slice"):
                    k = int(s.attrib["id"])
                    if 0 <= k < PROFILE_SIZE:
                        p[k]["count_equiv"] = float(s.attrib.get("countEquiv", s.attrib.get("count", 0)))
                        p[k]["sum_1_per_ort_velocity"] = float(s.attrib.get("sum_1_per_v", 0))
                        p[k]["sum_v_ort"] = float(s.attrib.get("sum_v_ort", 0))
            tx = fres.find("Texture")
            if tx is not None and off["texture"][f] >= 0:
                w, hgt = int(tx.attrib["width"]), int(tx.attrib["height"])
                cells = t.textures[m][off["texture"][f]:off["texture"][f] + w * hgt]
                cnode = tx.find("countEquiv") if tx.find("countEquiv") is not None else tx.find("count")
                for name, node in (("count_equiv", cnode), ("sum_1_per_ort_velocity", tx.find("sum_1_per_v")),
                                   ("sum_v_ort_per_area", tx.find("sum_v_ort"))):
                    if node is not None:
                        v = _numbers(node)
                        cells[name][: min(len(v), len(cells))] = v[: len(cells)]
            dn = fres.find("Directions")
            if dn is not None and off["direction"][f] >= 0:
                n = int(dn.attrib["width"]) * int(dn.attrib["height"])
                cells = t.directions[m][off["direction"][f]:off["direction"][f] + n]
                vv = _numbers(dn.find("vel.vectors")).reshape(-1, 3) if dn.find("vel.vectors") is not None else None
                cc = _numbers(dn.find("count")) if dn.find("count") is not None else None
                if vv is not None:
                    cells["dir"][: len(vv)] = vv[: len(cells)]
                if cc is not None:
                    cells["count"][: len(cc)] = cc[: len(cells)].astype(np.uint64)
            if off["facet_hist"][f] >= 0:
                sz = facet_hist_sizes(F[f])
                _hist_into(fres.find("Histograms"), t.facet_hist[m][off["facet_hist"][f]:], sz)
    return t, des_limit


def _angle_maps_from_tree(root: ET.Element, model: Model, t) -> None:
    fnodes = root.find("Geometry").find("Facets").findall("Facet")
    off = t.offsets["anglemap"]
    F = model.facets
    for f, fn in enumerate(fnodes):
        if off[f] < 0:
            continue
        ia = fn.find("IncidentAngleMap")
        if ia is None or ia.find("map") is None:
            continue
        n = int(F["am_phi_width"][f]) * int(F["am_theta_lower_res"][f] + F["am_theta_higher_res"][f])
        vals = np.array([int(float(x)) for x in (ia.find("map").text or "").split()], dtype=np.uint64)
        t.angle_maps[off[f]:off[f] + min(n, len(vals))] = vals[:n]


# ---------------------------------------------------------------- writing
def _cdata_block(rows) -> str:
    return "\n" + "".join("\t".join(r) + "\t\n" for r in rows)


def _write_back_parameters(root: ET.Element, model: Model) -> None:
    """Store the parameters that ad hoc parameter changes may alter
    (sticking, opacity, outgassing and its desorption type, temperature,
    gas mass, decay) in the geometry part, as the reference's CLI does
    before writing results (src/MolflowCLI.cpp:273-305). Attributes whose
    value is unchanged are left as they are."""
    def put(node, attr, value, as_int=False):
        if node is None:
            return
        old = node.attrib.get(attr)
        try:
            same = old is not None and (int(float(old)) == int(value) if as_int else float(old) == float(value))
        except ValueError:
            same = False
        if not same:
            node.set(attr, str(int(value)) if as_int else repr(float(value)))

    F = model.facets
    fnodes = root.find("Geometry").find("Facets").findall("Facet")
    for f, fn in enumerate(fnodes):
        put(fn.find("Sticking"), "constValue", F["sticking"][f])
        put(fn.find("Opacity"), "constValue", F["opacity"][f])
        put(fn.find("Outgassing"), "constValue", F["outgassing"][f])
        put(fn.find("Outgassing"), "desType", F["desorb_type"][f], as_int=True)
        put(fn.find("Temperature"), "value", F["temperature"][f])
    ss = root.find("MolflowSimuSettings")
    gas = ss.find("Gas") if ss is not None else None
    s = model.settings
    put(gas, "mass", s.gas_mass)
    put(gas, "halfLife", s.half_life)
    if gas is not None and ("enableDecay" in gas.attrib or s.enable_decay):
        put(gas, "enableDecay", s.enable_decay, as_int=True)


def save(path: str, doc: XmlDocument, tallies, des_limit: int = 0, molecules_per_tp=None) -> None:
    """Write `doc` with its results replaced by `tallies` (the layout of
    SaveSimulationState, WriterXML.cpp:367-677) and the facets' recorded
    angle maps updated. molecules_per_tp(m), if given, fills the
    informational molecules_per_TP attribute of each moment."""
    root = doc.root
    model = doc.model
    F = model.facets
    t = tallies
    off = t.offsets
    _write_back_parameters(root, model)
    # recorded angle maps are stored with the facets
    fnodes = root.find("Geometry").find("Facets").findall("Facet")
    for f, fn in enumerate(fnodes):
        if off["anglemap"][f] < 0:
            continue
        W, T = int(F["am_phi_width"][f]), int(F["am_theta_lower_res"][f] + F["am_theta_higher_res"][f])
        vals = t.angle_maps[off["anglemap"][f]:off["anglemap"][f] + W * T].reshape(T, W)
        ia = fn.find("IncidentAngleMap")
        if ia is None:
            ia = ET.SubElement(fn, "IncidentAngleMap")
        ia.attrib.update(angleMapPhiWidth=str(W), angleMapThetaLimit=repr(float(F["am_theta_limit"][f])),
                         angleMapThetaLowerRes=str(int(F["am_theta_lower_res"][f])),
                         angleMapThetaHigherRes=str(int(F["am_theta_higher_res"][f])))
        mp = ia.find("map")
        if mp is None:
            mp = ET.SubElement(ia, "map")
        mp.text = _cdata_block([[str(int(x)) for x in row] for row in vals])
    old = root.find("MolflowResults")
    if old is not None:
        root.remove(old)
    res = ET.SubElement(root, "MolflowResults")
    moms = ET.SubElement(res, "Moments", nb=str(t.n_slots))
    g = t.glob[0]
    gs = model.settings
    gsz = global_hist_sizes(gs)
    moments = sorted(model.moments, key=lambda mw: mw[0] - 0.5 * mw[1])
    for m in range(t.n_slots):
        mom = ET.SubElement(moms, "Moment", id=str(m))
        if m == 0:
            mom.set("time", "Constant flow")
            mom.set("timeWindow", "0")
        else:
            mom.set("time", repr(float(moments[m - 1][0])))
            mom.set("timeWindow", repr(float(moments[m - 1][1])))
        mom.set("molecules_per_TP", repr(float(molecules_per_tp(m))) if molecules_per_tp else "0")
        if m == 0:
            gl = ET.SubElement(mom, "Global")
            ET.SubElement(gl, "Hits", totalHit=str(int(g["nb_mc_hit"])), totalHitEquiv=repr(float(g["nb_hit_equiv"])),
                          totalDes=str(int(g["nb_desorbed"])), totalAbsEquiv=repr(float(g["nb_abs_equiv"])),
                          totalDist_total=repr(float(g["dist_traveled_total"])),
                          totalDist_fullHitsOnly=repr(float(g["dist_traveled_full_hits_only"])),
                          totalLeak=str(int(g["nb_leak_total"])), maxDesorption=str(int(des_limit)))
            ET.SubElement(gl, "Hit_Cache", nb="0")
            nl = int(min(t.leak_count[0], len(t.leaks)))
            lc = ET.SubElement(gl, "Leak_Cache", nb=str(nl))
            for i in range(nl):
                p = t.leaks[i]
                ET.SubElement(lc, "Leak", id=str(i), posX=repr(p[0]), posY=repr(p[1]), posZ=repr(p[2]),
                              dirX=repr(p[3]), dirY=repr(p[4]), dirZ=repr(p[5]))
        if sum(gsz):
            _hist_write(ET.SubElement(mom, "Histograms"), t.global_hist[m],
                        (gs.gh_record_bounce, gs.gh_bounce_binsize, gs.gh_bounce_max),
                        (gs.gh_record_distance, gs.gh_distance_binsize, gs.gh_distance_max),
                        (gs.gh_record_time, gs.gh_time_binsize, gs.gh_time_max), gsz)
        frs = ET.SubElement(mom, "FacetResults")
        for f in range(model.n_facets):
            fe = ET.SubElement(frs, "Facet", id=str(f))
            c = t.facets[m][f]
            ET.SubElement(fe, "Hits", nbHit=str(int(c["nb_mc_hit"])), nbHitEquiv=repr(float(c["nb_hit_equiv"])),
                          nbDes=str(int(c["nb_desorbed"])), nbAbsEquiv=repr(float(c["nb_abs_equiv"])),
                          sum_v_ort=repr(float(c["sum_v_ort"])), sum_1_per_v=repr(float(c["sum_1_per_ort_velocity"])),
                          sum_v=repr(float(c["sum_1_per_velocity"])))
            if gs.enable_force:
                fo = ET.SubElement(fe, "Forces")
                for k, tag in enumerate(("Impulse", "Impulse_square", "Impulse_momentum")):
                    v = t.forces[m][f][3 * k:3 * k + 3]
                    ET.SubElement(fo, tag, x=repr(float(v[0])), y=repr(float(v[1])), z=repr(float(v[2])))
            if off["profile"][f] >= 0:
                pe = ET.SubElement(fe, "Profile", size=str(PROFILE_SIZE))
                p = t.profiles[m][off["profile"][f]:off["profile"][f] + PROFILE_SIZE]
                for k in range(PROFILE_SIZE):
                    ET.SubElement(pe, "Slice", id=str(k), countEquiv=repr(float(p[k]["count_equiv"])),
                                  sum_1_per_v=repr(float(p[k]["sum_1_per_ort_velocity"])),
                                  sum_v_ort=repr(float(p[k]["sum_v_ort"])))
            if off["texture"][f] >= 0:
                w, h = texture_cells(F["tex_width_precise"][f]), texture_cells(F["tex_height_precise"][f])
                cells = t.textures[m][off["texture"][f]:off["texture"][f] + w * h].reshape(h, w)
                te = ET.SubElement(fe, "Texture", width=str(w), height=str(h))
                for tag, name in (("count", "count_equiv"), ("sum_1_per_v", "sum_1_per_ort_velocity"),
                                  ("sum_v_ort", "sum_v_ort_per_area")):
                    ET.SubElement(te, tag).text = _cdata_block([[repr(float(x)) for x in row] for row in cells[name]])
            if off["direction"][f] >= 0:
                w, h = texture_cells(F["tex_width_precise"][f]), texture_cells(F["tex_height_precise"][f])
                cells = t.directions[m][off["direction"][f]:off["direction"][f] + w * h].reshape(h, w)
                de = ET.SubElement(fe, "Directions", width=str(w), height=str(h))
                ET.SubElement(de, "vel.vectors").text = _cdata_block(
                    [[",".join(repr(float(x)) for x in cell) for cell in row["dir"]] for row in cells])
                ET.SubElement(de, "count").text = _cdata_block([[str(int(x)) for x in row["count"]] for row in cells])
            if off["facet_hist"][f] >= 0:
                sz = facet_hist_sizes(F[f])
                _hist_write(ET.SubElement(fe, "Histograms"), t.facet_hist[m][off["facet_hist"][f]:],
                            (F["fh_record_bounce"][f], F["fh_bounce_binsize"][f], F["fh_bounce_max"][f]),
                            (F["fh_record_distance"][f], F["fh_distance_binsize"][f], F["fh_distance_max"][f]),
                            (F["fh_record_time"][f], F["fh_time_binsize"][f], F["fh_time_max"][f]), sz)
    tree = ET.ElementTree(root)
    ET.indent(tree, space="  ")
    tmp = path + ".tmp"
    tree.write(tmp, encoding="utf-8", xml_declaration=True)
    os.replace(tmp, path)


def _hist_write(node, h, bounce, dist, time_, sizes):
    o = 0
    for tag, (rec, binsize, mx), n in zip(("Bounces", "Distance", "Time"), (bounce, dist, time_), sizes):
        if not n:
            continue
        e = ET.SubElement(node, tag, size=str(n), binSize=repr(float(binsize)) if tag != "Bounces" else str(int(binsize)),
                          max=repr(float(mx)) if tag != "Bounces" else str(int(mx)))
        for k in range(n):
            start = "overRange" if k == n - 1 else repr(float(k * binsize)) if tag != "Bounces" else str(int(k * binsize))
            ET.SubElement(e, "Bin", start=start, count=repr(float(h[o + k])))
        o += n
