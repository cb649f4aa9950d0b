"""molflowCLI on the GPU (molflow_sources_amd/cli.py), on two of the
reference's test case files (tests/golden/xml, copied from the reference's
tests/TestCases): runs to a desorption limit, continues from a result file,
stops on a time limit with every particle complete, applies parameter
changes and writes the facet exports. Needs an MI355X."""
import math
import os

import numpy as np
import pytest

from molflow_sources_amd import xml_io
from molflow_sources_amd.compare import TallyState, compare_states

pytestmark = pytest.mark.gpu

XML = os.path.join(os.path.dirname(__file__), "golden", "xml")
B01 = os.path.join(XML, "B01-lr1000_pipe.zip")
Q01 = os.path.join(XML, "01-quick_pipe_profiles_textures_2sided.zip")
QUIET = ["-s", "0", "-a", "0", "--noProgress"]


def cli(*args):
    from molflow_sources_amd import cli as C
    return C.main([str(a) for a in args])


def state(doc):
    return TallyState(doc.results, doc.model)


def test_reset_run_passes_compare_with_the_stored_results(tmp_path):
    """The reference's ResultsOkay protocol through the CLI: --reset, run
    to the stored desorption count (716,758,831 on test case 01), then
    Compare (thresholds 0.009 / 0.07) against the results the input file
    stores. The count must match because Compare checks the direction
    cells' raw sums."""
    stored = xml_io.load(Q01)
    n = int(stored.results.glob[0]["nb_desorbed"])
    assert cli("-f", Q01, "--reset", "-d", n, "--desPerStep", 100_000_000, "--outputPath", tmp_path,
               "-o", "r.xml", *QUIET) == 0
    out = xml_io.load(str(tmp_path / "r.xml"))
    assert int(out.results.glob[0]["nb_desorbed"]) == n
    assert out.des_limit == n
    cnt = {}
    g, f, fine, msgs = compare_states(state(out), state(stored), glob=0.009, loc=0.07, counts=cnt)
    assert g == 0 and f == 0, msgs[:5]
    assert fine <= 0.02 * cnt["fine_checks"] + 2, (fine, cnt, msgs[:5])


def test_reset_run_transmission_of_the_long_pipe(tmp_path):
    """Test case B01 (L/R = 1000): the number of hits per particle is too
    heavy-tailed for a 0.9 % global check at these counts, so the check is
    the transmission (outlet absorptions / desorptions), a binomial
    proportion, against the stored one."""
    n = 400_000
    assert cli("-f", B01, "--reset", "-d", n, "--outputPath", tmp_path, "-o", "r.xml", *QUIET) == 0
    out = xml_io.load(str(tmp_path / "r.xml")).results
    st = xml_io.load(B01).results
    assert int(out.glob[0]["nb_desorbed"]) == n
    p1, n1 = out.facets[0]["nb_abs_equiv"][1] / n, n
    n2 = int(st.glob[0]["nb_desorbed"])
    p2 = st.facets[0]["nb_abs_equiv"][1] / n2
    p = (p1 * n1 + p2 * n2) / (n1 + n2)
    z = (p1 - p2) / math.sqrt(p * (1 - p) * (1 / n1 + 1 / n2))
    assert abs(z) < 4, (p1, p2, z)


def test_continuing_a_result_file_equals_one_longer_run(tmp_path):
    """A run continued from its own result file takes the particles after
    the ones already run: 100k + 50k desorptions give the same integer
    counters as one run of 150k."""
    a, b = tmp_path / "a", tmp_path / "b"
    assert cli("-f", B01, "--reset", "-d", 100_000, "--outputPath", a, "-o", "r.xml", *QUIET) == 0
    assert cli("-f", a / "r.xml", "-d", 150_000, "--outputPath", a, "-o", "r2.xml", *QUIET) == 0
    assert cli("-f", B01, "--reset", "-d", 150_000, "--outputPath", b, "-o", "r.xml", *QUIET) == 0
    two = xml_io.load(str(a / "r2.xml")).results
    one = xml_io.load(str(b / "r.xml")).results
    for k in ("nb_mc_hit", "nb_desorbed", "nb_leak_total"):
        assert one.glob[0][k] == two.glob[0][k], k
    assert np.array_equal(one.facets["nb_mc_hit"], two.facets["nb_mc_hit"])
    assert np.array_equal(one.facets["nb_desorbed"], two.facets["nb_desorbed"])
    assert np.allclose(one.facets["nb_abs_equiv"], two.facets["nb_abs_equiv"], rtol=0, atol=0)
    # a limit the file has already reached runs nothing
    assert cli("-f", b / "r.xml", "-d", 150_000, "--outputPath", b, "-o", "same.xml", *QUIET) == 0
    assert xml_io.load(str(b / "same.xml")).results.glob[0]["nb_desorbed"] == 150_000


def test_time_limit_finishes_every_particle(tmp_path):
    """-t 2: the run stops after about two seconds and the particles in
    flight are traced to their end, so every desorbed particle is absorbed
    (the pipe's walls do not stick, its two ends do)."""
    import time
    t0 = time.perf_counter()
    assert cli("-f", B01, "--reset", "-t", 2, "--desPerStep", 200_000, "--outputPath", tmp_path, "-o", "r.xml",
               *QUIET) == 0
    assert time.perf_counter() - t0 < 60
    r = xml_io.load(str(tmp_path / "r.xml")).results
    n_des = int(r.glob[0]["nb_desorbed"])
    assert n_des > 100_000
    assert r.glob[0]["nb_abs_equiv"] == n_des
    assert r.facets[0]["nb_abs_equiv"].sum() == n_des


def test_parameter_changes_and_exports(tmp_path):
    """--setParams: the outlet (facet 2) made half transparent lets half
    the particles that reach it leave the model; the changed value is
    written to the result file, and the two CSV exports are written with
    one row per facet."""
    base = tmp_path / "base"
    assert cli("-f", B01, "--reset", "-d", 200_000, "--outputPath", base, "-o", "r.xml", *QUIET) == 0
    ch = tmp_path / "ch"
    assert cli("-f", B01, "--reset", "-d", 200_000, "--setParams", "facet.2.opacity=0.5", "--outputPath", ch,
               "-o", "r.xml", "--writeFacetDetails", "--writeFacetQuantities", *QUIET) == 0
    d0 = xml_io.load(str(base / "r.xml"))
    d1 = xml_io.load(str(ch / "r.xml"))
    assert d1.model.facets["opacity"][1] == 0.5 and d0.model.facets["opacity"][1] == 1.0
    # half the particles reaching the outlet pass it and leave the model
    # (leaks), the rest are absorbed there
    g0, g1 = d0.results.glob[0], d1.results.glob[0]
    a0 = d0.results.facets[0]["nb_abs_equiv"][1]
    a1 = d1.results.facets[0]["nb_abs_equiv"][1]
    assert g0["nb_leak_total"] == 0 and g0["nb_abs_equiv"] == 200_000
    assert g1["nb_abs_equiv"] + g1["nb_leak_total"] == 200_000
    assert abs(g1["nb_leak_total"] - a1) < 5 * math.sqrt(a0), (g1["nb_leak_total"], a1, a0)
    for fn in ("facet_details.csv", "facet_physics.csv"):
        rows = (ch / fn).read_text().splitlines()
        assert len(rows) == d1.model.n_facets + 1, fn
    # invalid changes stop the run before it starts
    assert cli("-f", B01, "--reset", "-d", 1000, "--setParams", "facet.2.sticking=3", "--outputPath", ch,
               "-o", "bad.xml", *QUIET) == 42
    assert not (ch / "bad.xml").exists()


def test_zip_output(tmp_path):
    import zipfile
    assert cli("-f", B01, "--reset", "-d", 10_000, "--outputPath", tmp_path, "-o", "r.zip", *QUIET) == 0
    with zipfile.ZipFile(tmp_path / "r.zip") as z:
        assert z.namelist() == ["r.xml"]
    assert int(xml_io.load(str(tmp_path / "r.zip")).results.glob[0]["nb_desorbed"]) == 10_000
