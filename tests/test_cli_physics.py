"""The layers next to the tracer, on the CPU: ad hoc parameter changes
(molflow_sources_amd/params.py, the reference's ParameterParser), the
physical quantities of the facet exports (physics.py, CSVExporter) and the
command line driver's argument handling (cli.py).

The parameter tests restate the reference's own ParameterParsing tests
(tests/testsuite.cpp:1078-1210) on a 200-facet model. The physics tests pin
the export formulas with the kinetic theory of a gas in equilibrium, on
oracle tallies of a closed isothermal box: P = n k T and Z = n <v> / 4
(a dropped 0.01, 1e4, mass or 2 fails one of them)."""
import math
import os

import numpy as np
import pytest

import feature_models as FM
from molflow_sources_amd import _abi as A
from molflow_sources_amd import params as P
from molflow_sources_amd import physics
from molflow_sources_amd.model import Model, default_facets

PAM3S_TO_MBARLS = 10.0


def two_hundred_facets() -> Model:
    """200 unit right triangles; facet 91 (index 90) has area 123 cm^2."""
    V, polys = [], []
    for i in range(200):
        leg = math.sqrt(2 * 123.0) if i == 90 else math.sqrt(2.0)
        V += [[0, 0, i], [leg, 0, i], [0, leg, i]]
        polys.append([3 * i, 3 * i + 1, 3 * i + 2])
    return Model(np.array(V, dtype=np.float64), polys, default_facets(200))


def check_reference_changes(m: Model):
    F = m.facets
    assert F["opacity"][41] == 0.5
    assert F["sticking"][2] == 10.01
    for f in (49, 69, 89):
        assert F["outgassing"][f] * PAM3S_TO_MBARLS == pytest.approx(42e5, rel=1e-15)
        assert F["desorb_type"][f] == A.DES_COSINE          # outgassing switched on: cosine source
    assert F["outgassing"][48] == 0.0 and F["outgassing"][90 - 0] != 0.0 and F["outgassing"][91] == 0.0
    assert F["outgassing"][90] * PAM3S_TO_MBARLS == pytest.approx(1.23e-10, rel=1e-12)
    assert F["temperature"][99] == 290.92


def test_parameter_change_file(tmp_path):
    """ParameterParsing.ParameterChangeFile."""
    pf = tmp_path / "change.cfg"
    pf.write_text("facet.42.opacity=0.5\n"
                  "facet.3.sticking=10.01\n"
                  "facet.50-90.outgassing=42e5\n"
                  "facet.91.specificOutgassing=1e-12\n"
                  "facet.100.temperature=290.92\n"
                  "simulation.mass=42.42\n"
                  "simulation.enableDecay=1\n"
                  "simulation.halfLife=42.42")
    m = two_hundred_facets()
    assert m.settings.gas_mass != 42.42 and not m.settings.enable_decay
    ch = P.parse_file(str(pf))
    assert not ch.messages
    log = []
    n_err = P.apply(m, ch, log=log.append)
    assert m.settings.gas_mass == 42.42 and m.settings.enable_decay == 1 and m.settings.half_life == 42.42
    check_reference_changes(m)
    # the sticking of 10.01 is applied and reported as the one invalid value
    assert n_err == 1 and "Invalid sticking coefficient" in log[0]


def test_parameter_change_list():
    """ParameterParsing.ParameterChangeList: one change per argument, and
    the same changes joined with ';' in one argument."""
    args = ["facet.42.opacity=0.5", "facet.3.sticking=10.01", "facet.50-90.outgassing=42e5",
            "facet.91.specificOutgassing=1e-12", "facet.100.temperature=290.92", "simulation.mass=42.42"]
    for given in (args, [";".join(args)]):
        m = two_hundred_facets()
        P.apply(m, P.parse(given), log=lambda s: None)
        assert m.settings.gas_mass == 42.42
        check_reference_changes(m)


def test_parameter_groups():
    """ParameterParsing.Group: a quoted selection name stands for its
    facets (0-based ids in the file)."""
    sel = {"Valid Selection": [5, 8, 9], "InvalidSelection": [1, 2, 3, 4], "NotValid": [10, 11, 12]}
    ch = P.parse(['facet."NotValid".opacity=0.3', 'facet."Valid Selection".opacity=0.5',
                  'facet."InvalidSelection".opacity=0.8'], sel)
    m = two_hundred_facets()
    P.apply(m, ch, log=lambda s: None)
    op = m.facets["opacity"]
    assert op[5] == op[8] == op[9] == 0.5
    assert op[4] == 0.8 and op[10] == 0.3 and op[6] == op[7] == 1.0


def test_parameter_errors_are_reported():
    ch = P.parse(["facet.1.colour=3", "simulation.speed=1", "gas.mass=3", 'facet."nope".opacity=1',
                  "facet.0.opacity=1", "facet.5-2.opacity=1"], {})
    assert not ch.facet and not ch.simulation
    assert len(ch.messages) == 6
    m = two_hundred_facets()
    # facet ids beyond the model and out-of-range opacities count as errors
    log = []
    assert P.apply(m, P.parse(["facet.201.opacity=0.5", "facet.1.opacity=1.5"]), log=log.append) == 2
    assert m.facets["opacity"][0] == 1.5
    # values are read like strtod: an unreadable value is 0
    P.apply(m, P.parse(["facet.2.sticking=abc"]), log=log.append)
    assert m.facets["sticking"][1] == 0.0


def test_selections_read_from_the_interface_block():
    import xml.etree.ElementTree as ET
    root = ET.fromstring('<SimulationEnvironment><Interface><Selections nb="2">'
                         '<Selection id="0" name="pumps"><selItem id="0" facet="3"/><selItem id="1" facet="7"/>'
                         '</Selection><Selection id="1" name="walls"><selItem id="0" facet="0"/></Selection>'
                         '</Selections></Interface></SimulationEnvironment>')
    assert P.selections_of(root) == {"pumps": [3, 7], "walls": [0]}


# ---------------------------------------------------------------- physics
@pytest.fixture(scope="module")
def equilibrium(oracle_lib):
    from oracle.mfcpu import CpuModel
    m = FM.equilibrium_box()
    cm = CpuModel(m)
    t = cm.run(3000, seed=11, threads=8)
    t.offsets = cm.offsets()
    return m, t, physics.model_totals(m), 293.15


def test_model_totals_closed_form():
    """One source of Q = 1 Pa m3/s at 293.15 K: Q / (k T) molecules per
    second (k = 1.38e-23 as in the reference), Q in Pa m3/s, and the total
    desorbed up to the latest moment (half the time window without moments)
    is that rate times the latest moment."""
    m = FM.equilibrium_box()
    tot = physics.model_totals(m)
    rate = 1.0 / (FM.KB * 293.15)
    assert tot.final_outgassing_rate == pytest.approx(rate, rel=1e-15)
    assert tot.final_outgassing_rate_pa_m3_s == 1.0
    assert tot.latest_moment == pytest.approx(0.5 * m.settings.time_window, rel=1e-15)
    assert tot.total_desorbed_molecules == pytest.approx(rate * tot.latest_moment, rel=1e-15)


def test_model_totals_time_dependent_match_the_oracle(oracle_lib):
    """Time-dependent outgassing (test case 02): the integrated desorption
    and the final rate of the library and the oracle agree."""
    from fixture_models import model_from_fixture
    from oracle.mfcpu import CpuModel
    m, _ = model_from_fixture("02-timedependent_with_two_parameters")
    tot = physics.model_totals(m)
    cm = CpuModel(m)
    assert tot.total_desorbed_molecules == pytest.approx(cm.total_desorbed_molecules, rel=1e-14)
    assert tot.final_outgassing_rate == pytest.approx(cm.final_outgassing_rate, rel=1e-14)
    assert tot.latest_moment == cm.latest_moment


WALLS = ("zlo", "zhi", "xlo", "xhi", "ylo", "yhi")


def test_ideal_gas_law_on_the_walls(equilibrium):
    """Pressure and density of the export formulas obey P = n k T on walls
    in equilibrium, k T / m being the speed scale of the tracer's Maxwell
    sampling (1.38e-23 T / (m 1.67e-27)) and m the formulas' molecule mass
    (m / 1000 / 6e23 kg)."""
    m, t, totals, T = equilibrium
    q = physics.facet_quantities(m, t, totals)
    ids = [m.ids[w] for w in WALLS]
    assert all(t.facets[0]["nb_mc_hit"][i] > 100_000 for i in ids)
    kT = FM.KB * T * (1e-3 / 6e23) / FM.AMU
    P_pa = q["pressure"][ids] * 100.0                     # mbar -> Pa
    ratio = P_pa / (q["density"][ids] * kT)
    assert np.allclose(ratio, 1.0, atol=0.02), ratio


def test_impingement_is_density_times_mean_speed_over_four(equilibrium):
    m, t, totals, T = equilibrium
    q = physics.facet_quantities(m, t, totals)
    ids = [m.ids[w] for w in WALLS]
    v_mean = math.sqrt(8 * FM.KB * T / (math.pi * FM.AMU * m.settings.gas_mass))
    assert np.allclose(q["avg_speed"][ids] / v_mean, 1.0, atol=0.01), q["avg_speed"][ids]
    z = q["density"][ids] * q["avg_speed"][ids] / 4.0
    assert np.allclose(q["impingement"][ids] / z, 1.0, atol=0.02)


def test_steady_state_particle_balance(equilibrium):
    """Every desorbed particle ends on the pump: its absorption rate
    (absorptions x molecules per test particle) is the outgassing rate."""
    m, t, totals, T = equilibrium
    pump = m.ids["pump"]
    n_des = int(t.glob[0]["nb_desorbed"])
    mptp = physics.molecules_per_tp(0, totals, n_des, m)
    assert t.facets[0]["nb_abs_equiv"][pump] * mptp == pytest.approx(totals.final_outgassing_rate, rel=1e-12)
    # the pump absorbs and the source desorbs: their densities are corrected
    dc = physics.density_correction(t.facets[0])
    assert dc[m.ids["zlo"]] == 1.0 and 0.5 <= dc[pump] < 1.0 and dc[m.ids["source"]] < 1.0


def test_formulas_over_the_variables(equilibrium):
    m, t, totals, T = equilibrium
    v = physics.formula_variables(m, t, totals)
    pump = m.ids["pump"] + 1
    assert physics.evaluate_formula(f"A{pump}/SUMDES", v) == 1.0
    assert physics.evaluate_formula("SUMDES", v) == 3000
    assert physics.evaluate_formula(f"ar{pump} * 2", v) == pytest.approx(2.0)
    assert physics.evaluate_formula("sqrt(4) + 2^3 - PI", v) == pytest.approx(10.0 - math.pi)
    with pytest.raises(KeyError):
        physics.evaluate_formula("Q9999", v)
    with pytest.raises(ValueError):
        physics.evaluate_formula("__import__('os')", v)


def test_csv_exports(equilibrium, tmp_path):
    m, t, totals, T = equilibrium
    physics.write_facet_quantities_csv(str(tmp_path / "facet_physics.csv"), m, t, totals)
    physics.write_facet_details_csv(str(tmp_path / "facet_details.csv"), m, t, totals)
    rows = (tmp_path / "facet_physics.csv").read_text().splitlines()
    assert len(rows) == m.n_facets + 1 and rows[0].startswith("#,Imping.rate,Pressure [mbar]")
    q = physics.facet_quantities(m, t, totals)
    for f in range(m.n_facets):
        cells = rows[f + 1].split(",")
        assert float(cells[2]) == q["pressure"][f]
    det = (tmp_path / "facet_details.csv").read_text().splitlines()
    assert len(det) == m.n_facets + 1
    assert all(len(r.split(",")) == len(det[0].split(",")) for r in det)


# ---------------------------------------------------------------- CLI arguments
def test_cli_refuses_runs_without_a_stop_condition(tmp_path, capsys):
    from molflow_sources_amd import cli
    xml = os.path.join(os.path.dirname(__file__), "golden", "xml", "B01-lr1000_pipe.zip")
    assert cli.main(["-f", xml, "--outputPath", str(tmp_path)]) == 41
    assert "No stop condition" in capsys.readouterr().err
    assert cli.main(["-f", xml, "-d", "10", "--overwrite", "-o", "x.xml"]) == 41
    assert cli.main(["-f", str(tmp_path / "missing.xml"), "-d", "10"]) == 41
    assert cli.main(["--bogus"]) == 41


def test_cli_output_names(tmp_path):
    from molflow_sources_amd import cli
    a = cli.build_parser().parse_args(["-f", "dir/B01.zip", "-d", "5", "--outputPath", str(tmp_path / "o")])
    outdir, out = cli.output_paths(a)
    assert outdir == str(tmp_path / "o") and out == os.path.join(outdir, "out_B01.zip")
    a = cli.build_parser().parse_args(["-f", "B01.zip", "-d", "5", "--outputPath", str(tmp_path / "o"),
                                       "-o", "r.xml"])
    assert cli.output_paths(a)[1] == os.path.join(str(tmp_path / "o"), "r.xml")
