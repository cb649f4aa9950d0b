"""The MolFlow XML loader and writer (molflow_sources_amd/xml_io.py).

Every test case file of the reference (tests/TestCases/*.zip, read where
the reference checkout is present) must load into the same model and the
same stored results as the committed fixture conversion
(tests/golden/*.json.gz); a written result file must load back unchanged.
No GPU needed."""
import math
import os

import numpy as np
import pytest

from fixture_models import ALL_FIXTURES, StoredState, load_json, model_from_fixture
from molflow_sources_amd import xml_io

CASES = "/root/reference/tests/TestCases"
CATALOG = "/root/reference/copy_to_build/parameter_catalog"


COMMITTED = os.path.join(os.path.dirname(__file__), "golden", "xml")


def fixture_zip(name):
    """The test case file: one of the two copies committed under
    tests/golden/xml, else the reference checkout's."""
    for d in (COMMITTED, CASES):
        if not os.path.isdir(d):
            continue
        for fn in os.listdir(d):
            if fn.endswith(".zip") and fn[:-4].replace(" ", "_") == name:
                return os.path.join(d, fn)
    return None


needs_reference = pytest.mark.skipif(not os.path.isdir(CASES), reason="reference test cases not present")


@pytest.mark.parametrize("name", ["B01-lr1000_pipe", "01-quick_pipe_profiles_textures_2sided"])
def test_committed_test_cases_load(name):
    """The two test case files kept in the repository load (without the
    parameter catalog, which neither uses) into the converted fixture's
    model and stored totals."""
    doc = xml_io.load(os.path.join(COMMITTED, name + ".zip"))
    ref, d = model_from_fixture(name)
    assert np.array_equal(doc.model.vertices, ref.vertices)
    for k in FIELDS:
        assert np.array_equal(doc.model.facets[k], ref.facets[k]), k
    S = StoredState(d)
    assert doc.results.glob[0]["nb_desorbed"] == S.glob[0]["nb_desorbed"]
    assert np.array_equal(doc.results.facets["nb_mc_hit"], S.facets["nb_mc_hit"])

FIELDS = ["sticking", "opacity", "temperature", "accomodation", "outgassing", "desorb_exponent", "diffuse_part",
          "specular_part", "cosine_exponent", "sojourn_freq", "sojourn_e", "tex_width_precise", "tex_height_precise",
          "am_theta_limit", "sticking_param", "opacity_param", "outgassing_param", "temperature_param", "desorb_type",
          "is_2sided", "super_idx", "super_dest", "teleport_dest", "profile_type", "is_moving", "enable_sojourn",
          "is_textured", "count_des", "count_abs", "count_refl", "count_trans", "count_dir", "am_record",
          "am_phi_width", "am_theta_lower_res", "am_theta_higher_res", "fh_record_bounce", "fh_bounce_binsize",
          "fh_bounce_max", "fh_record_distance", "fh_distance_binsize", "fh_distance_max", "fh_record_time",
          "fh_time_binsize", "fh_time_max"]


@needs_reference
@pytest.mark.parametrize("name", ALL_FIXTURES)
def test_loader_reproduces_the_converted_fixture(name):
    path = fixture_zip(name)
    assert path, name
    doc = xml_io.load(path, catalog_dir=CATALOG)
    ref, d = model_from_fixture(name)
    m = doc.model
    assert np.array_equal(m.vertices, ref.vertices)
    assert m.polygons == ref.polygons
    assert m.n_structures == ref.n_structures
    for k in FIELDS:
        assert np.array_equal(m.facets[k], ref.facets[k]), k
    for k, _ in m.settings._fields_:
        a, b = getattr(m.settings, k), getattr(ref.settings, k)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), k
        else:
            assert a == b, k
    assert [p.name for p in m.params] == [p.name for p in ref.params]
    for p, q in zip(m.params, ref.params):
        assert list(p.x) == list(q.x) and list(p.y) == list(q.y) and p.log_x == q.log_x and p.log_y == q.log_y
    assert m.moments == ref.moments
    assert sorted(m.am_maps) == sorted(ref.am_maps)
    for f in m.am_maps:
        assert np.array_equal(m.am_maps[f], ref.am_maps[f])
    # stored results: global and per-slot facet counters, stored profiles and textures
    S = StoredState(d)
    t = doc.results
    assert t is not None
    for k in ("nb_mc_hit", "nb_desorbed", "nb_abs_equiv", "nb_leak_total"):
        assert t.glob[0][k] == S.glob[0][k], k
    assert t.facets.shape == S.facets.shape
    for k in ("nb_mc_hit", "nb_hit_equiv", "nb_desorbed", "nb_abs_equiv", "sum_v_ort", "sum_1_per_ort_velocity",
              "sum_1_per_velocity"):
        assert np.array_equal(t.facets[k], S.facets[k]), k
    from molflow_sources_amd.compare import TallyState
    ts = TallyState(t, m)
    for fid, prof in d["results"]["profiles"].items():
        for slot, rows in enumerate(prof):
            got = ts.profile(int(fid), slot)
            assert np.array_equal(got[0], np.asarray(rows)[:, 0])
    for fid, tex in d["results"]["textures"].items():
        got = ts.texture(int(fid), 0)
        assert np.array_equal(got[0], np.asarray(tex["count"]))
    for fid, h in d["results"]["facet_hist"].items():
        got = ts.facet_hist(int(fid), 0)
        for arr, k in zip(got, ("Bounces", "Distance", "Time")):
            if k in h:
                assert np.array_equal(arr, np.asarray(h[k])), k


@needs_reference
@pytest.mark.parametrize("name", ["01-quick_pipe_profiles_textures_2sided", "02-timedependent_with_two_parameters",
                                  "03-anglemap_record", "09-histograms"])
def test_written_results_load_back_unchanged(tmp_path, name):
    doc = xml_io.load(fixture_zip(name), catalog_dir=CATALOG)
    t = doc.results
    out = tmp_path / "result.xml"
    xml_io.save(str(out), doc, t, des_limit=1234)
    back = xml_io.load(str(out), catalog_dir=CATALOG)
    assert back.des_limit == 1234
    assert np.array_equal(back.model.vertices, doc.model.vertices)
    for k in FIELDS:
        assert np.array_equal(back.model.facets[k], doc.model.facets[k]), k
    b = back.results
    assert np.array_equal(b.glob, t.glob)
    assert np.array_equal(b.facets, t.facets)
    for name_ in ("textures", "profiles", "directions"):
        for fld in getattr(t, name_).dtype.names:
            assert np.array_equal(getattr(b, name_)[fld], getattr(t, name_)[fld]), (name_, fld)
    assert np.array_equal(b.angle_maps, t.angle_maps)
    assert np.array_equal(b.global_hist, t.global_hist)
    assert np.array_equal(b.facet_hist, t.facet_hist)


def test_loader_reads_outgassing_maps_and_writes_them_through(tmp_path):
    """A facet with a DynamicOutgassing map (LoaderXML.cpp:993-1037): width,
    height, ratios, total and the cells reach the model."""
    import xml.etree.ElementTree as ET
    root = ET.Element("SimulationEnvironment")
    geo = ET.SubElement(root, "Geometry")
    vs = ET.SubElement(geo, "Vertices")
    for i, (x, y, z) in enumerate([(0, 0, 0), (4, 0, 0), (4, 3, 0), (0, 3, 0)]):
        ET.SubElement(vs, "Vertex", id=str(i), x=str(x), y=str(y), z=str(z))
    ET.SubElement(geo, "Structures", nb="1")
    fs = ET.SubElement(geo, "Facets", nb="1")
    fn = ET.SubElement(fs, "Facet", id="0")
    ET.SubElement(fn, "Sticking", constValue="0")
    ET.SubElement(fn, "Opacity", constValue="1", is2sided="0")
    ET.SubElement(fn, "Outgassing", constValue="0", desType="2", desExponent="0", useOutgassingFile="1")
    ET.SubElement(fn, "Temperature", value="293.15", accFactor="1")
    ET.SubElement(fn, "Reflection", diffusePart="1", specularPart="0", cosineExponent="0")
    ET.SubElement(fn, "Structure", inStructure="0", linksTo="0")
    rec = ET.SubElement(fn, "Recordings")
    ET.SubElement(rec, "Profile", type="0")
    ET.SubElement(rec, "Texture", hasMesh="false")
    ET.SubElement(rec, "IncidentAngleMap", record="false")
    og = ET.SubElement(fn, "DynamicOutgassing", width="3", height="2", ratioU="0.75", ratioV="0.6",
                       totalDose="0", totalOutgassing="2.1e-3", totalFlux="0")
    ET.SubElement(og, "map").text = "\n1e-4\t2e-4\t3e-4\n4e-4\t5e-4\t6e-4\n"
    ind = ET.SubElement(fn, "Indices", nb="4")
    for i in range(4):
        ET.SubElement(ind, "Indice", i=str(i), vertex=str(i))
    ss = ET.SubElement(root, "MolflowSimuSettings")
    ET.SubElement(ss, "Gas", mass="28", halfLife="1e100")
    ET.SubElement(ss, "TimeSettings", timeWindow="1e-10", useMaxwellDistr="true", calcConstFlow="true")
    ET.SubElement(ss, "Motion", type="0")
    ET.SubElement(ss, "Parameters", nb="0")
    p = tmp_path / "og.xml"
    ET.ElementTree(root).write(str(p))
    doc = xml_io.load(str(p))
    m = doc.model
    assert m.facets["use_outgassing_map"][0] == 1
    assert (m.facets["og_map_width"][0], m.facets["og_map_height"][0]) == (3, 2)
    assert m.facets["og_ratio_u"][0] == 0.75 and m.facets["og_ratio_v"][0] == 0.6
    assert m.facets["total_outgassing"][0] == 2.1e-3
    assert np.allclose(m.og_maps[0], [[1e-4, 2e-4, 3e-4], [4e-4, 5e-4, 6e-4]])
    assert doc.results is None


def test_layout_matches_the_oracle(oracle_lib):
    """xml_io.tally_layout restates the library's tally layout rules; the
    oracle follows the same rules."""
    from oracle.mfcpu import CpuModel
    for name in ("01-quick_pipe_profiles_textures_2sided", "03-anglemap_record", "09-histograms",
                 "02-timedependent_with_two_parameters", "05-three_structures_nonsquare_textures",
                 "02c-timedependent_temperature"):
        m, _ = model_from_fixture(name)
        L, off = xml_io.tally_layout(m)
        cm = CpuModel(m)
        Lc = cm.layout()
        for k in ("n_facets", "n_slots", "n_tex_cells", "n_dir_cells", "n_profile_slices", "n_angle_cells",
                  "n_global_hist_bins", "n_facet_hist_bins"):
            assert getattr(L, k) == getattr(Lc, k), (name, k)
        oc = cm.offsets()
        for k in off:
            assert np.array_equal(off[k], oc[k]), (name, k)
