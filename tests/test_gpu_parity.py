"""Parity of the HIP tracer (through the C ABI) with the CPU oracle and with
the results stored in the reference's fixtures. Needs an MI355X."""
import math

import numpy as np
import pytest

from fixture_models import StoredState, model_from_fixture
from helpers import canonical_trace, polygon_zoo_model, random_rays, z_two_proportions
from molflow_sources_amd import geometry
from molflow_sources_amd.compare import TallyState, absorption_chi2, compare_states, results_wrong_detection
from oracle.mfcpu import CpuModel

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000


@pytest.fixture(scope="module")
def gpu():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.close()


def loaded(gpu, m, seed=42):
    gpu.load(m)
    gpu.rebuild_accel()
    gpu.reset(seed=seed)
    return gpu


# ---------------------------------------------------------------- intersection
INTERSECT_CASES = [
    ("tube", lambda: geometry.tube_model(), ([-1.2, -1.2, -0.5], [1.2, 1.2, 10.5])),
    ("beam_pipe", lambda: geometry.beam_pipe_model(), ([-5, -5, -1], [5, 5, 201])),
    ("quick_pipe_2sided", lambda: model_from_fixture("01-quick_pipe_profiles_textures_2sided")[0], ([-1, -1, 0], [1, 1, 5])),
    ("transparent_B04", lambda: model_from_fixture("B04-lr10_pipe_trans")[0], None),
]


@pytest.mark.parametrize("name,make,box", INTERSECT_CASES)
def test_first_hit_bit_exact(gpu, name, make, box):
    m = make()
    if box is None:
        lo, hi = m.vertices.min(0), m.vertices.max(0)
        box = (lo - 0.05 * (hi - lo), hi + 0.05 * (hi - lo))
    n = 50_000
    o, d = random_rays(n, box[0], box[1], seed=7)
    rng = np.random.default_rng(8)
    last = rng.integers(-1, m.n_facets, n).astype(np.int32)
    loaded(gpu, m)
    fg, tg, ug, vg = gpu.intersect(o, d, last)
    fc, tc, uc, vc = CpuModel(m).intersect(o, d, last, brute=True)
    assert np.array_equal(fg, fc), f"{int((fg != fc).sum())} first-hit facets differ"
    hit = fc >= 0
    assert hit.mean() > 0.3
    rel = np.abs(tg[hit] - tc[hit]) / tc[hit]
    assert rel.max() <= 1e-5
    assert np.allclose(ug[hit], uc[hit], atol=1e-9) and np.allclose(vg[hit], vc[hit], atol=1e-9)


def test_first_hit_bit_exact_4_wide(gpu):
    """The same fixed ray batch through a 4-wide build of the same format."""
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    o, d = random_rays(50_000, lo - 0.05 * (hi - lo), hi + 0.05 * (hi - lo), seed=17)
    gpu.load(m)
    gpu.rebuild_accel(4, 2)
    fg, tg, _, _ = gpu.intersect(o, d)
    fc, tc, _, _ = CpuModel(m).intersect(o, d, brute=True)
    gpu.rebuild_accel()
    assert np.array_equal(fg, fc), f"{int((fg != fc).sum())} first-hit facets differ"
    hit = fc >= 0
    assert np.all(np.abs(tg[hit] - tc[hit]) <= 1e-5 * tc[hit])


def test_first_hit_bit_exact_sixteen_facet_leaves(gpu):
    """The largest leaf a traversal stack entry holds (16 facets, a 4-bit
    count of 15) on the beam pipe: the same first hits as brute force; a
    leaf size of 17 is refused."""
    from molflow_sources_amd.simulation import MolflowGpuError
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    o, d = random_rays(30_000, [-3.9, -3.9, 1.0], [3.9, 3.9, 199.0], seed=37)
    gpu.load(m)
    gpu.rebuild_accel(8, 16)
    fg, tg, _, _ = gpu.intersect(o, d)
    fc, tc, _, _ = CpuModel(m).intersect(o, d, brute=True)
    assert np.array_equal(fg, fc), f"{int((fg != fc).sum())} first-hit facets differ"
    with pytest.raises(MolflowGpuError):
        gpu.rebuild_accel(8, 17)
    gpu.rebuild_accel()


@pytest.mark.parametrize("width,leaf", [(8, 1), (4, 1)])
def test_first_hit_bit_exact_one_facet_leaves(gpu, width, leaf):
    """Leaves of one facet (the shape the autotuner picks for the tube) on
    the tube and on the polygon set: the same first hits as brute force."""
    for m, lo, hi in ((geometry.tube_model(), [-1.2, -1.2, -0.5], [1.2, 1.2, 10.5]),
                      (polygon_zoo_model(), [-1.3, -1.3, -0.5], [1.3, 1.3, 3.5])):
        o, d = random_rays(30_000, lo, hi, seed=31)
        gpu.load(m)
        gpu.rebuild_accel(width, leaf)
        fg, tg, _, _ = gpu.intersect(o, d)
        fc, tc, _, _ = CpuModel(m).intersect(o, d, brute=True)
        assert np.array_equal(fg, fc), f"{m.name}: {int((fg != fc).sum())} first-hit facets differ"
        hit = fc >= 0
        assert np.all(np.abs(tg[hit] - tc[hit]) <= 1e-5 * tc[hit])


def test_first_hit_bit_exact_large_polygons(gpu):
    """Polygons of more than 4 vertices (non-convex star, comb with equal
    vertex heights and horizontal edges, 64-gon, C shape) go through the
    slab index of the point-in-polygon test: random rays plus axis-parallel
    rays on a grid, whose hit points sit at vertex heights."""
    m = polygon_zoo_model()
    o, d = random_rays(40_000, [-1.3, -1.3, -0.5], [1.3, 1.3, 3.5], seed=23)
    g = np.linspace(-1.0, 1.0, 201)
    X, Y = np.meshgrid(g, g)
    o2 = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 3.5)], 1)
    d2 = np.tile([0.0, 0.0, -1.0], (X.size, 1))
    o, d = np.concatenate([o, o2]), np.concatenate([d, d2])
    gpu.load(m)
    gpu.rebuild_accel()
    fg, tg, _, _ = gpu.intersect(o, d)
    fc, tc, _, _ = CpuModel(m).intersect(o, d, brute=True)
    assert np.array_equal(fg, fc), f"{int((fg != fc).sum())} first-hit facets differ"
    hit = fc >= 0
    assert hit.sum() > 20_000
    assert np.all(np.abs(tg[hit] - tc[hit]) <= 1e-5 * tc[hit])


def test_large_polygon_source_same_seed_tallies(gpu):
    """Desorption from a non-convex 22-vertex polygon samples the same
    positions as the oracle (rejection sampling through the slab index):
    the same seed gives the same integer tallies and leaks."""
    m = polygon_zoo_model()
    n = 200_000
    loaded(gpu, m, seed=13)
    gpu.run(n, max_events=MAXEV)
    g = gpu.read()
    c = CpuModel(m).run(n, seed=13, max_events=MAXEV)
    assert np.array_equal(g.facets["nb_mc_hit"].astype(np.int64), c.facets["nb_mc_hit"].astype(np.int64))
    assert g.glob[0]["nb_leak_total"] == c.glob[0]["nb_leak_total"]
    assert g.glob[0]["nb_desorbed"] == n


def test_first_hit_every_structure(gpu):
    m, _ = model_from_fixture("05-three_structures_nonsquare_textures")
    loaded(gpu, m)
    cm = CpuModel(m)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    o, d = random_rays(20_000, lo, hi, seed=9)
    for s in range(m.n_structures):
        fg, tg, _, _ = gpu.intersect(o, d, structure=s)
        fc, tc, _, _ = cm.intersect(o, d, structure=s, brute=True)
        assert np.array_equal(fg, fc)


# ---------------------------------------------------------------- event traces
TRACE_MODELS = ["tube", "01-quick_pipe_profiles_textures_2sided", "02-timedependent_with_two_parameters",
                "03b-anglemap_record_transparent_target", "05-three_structures_nonsquare_textures",
                "07-volume_decay", "08-wall_sojourn_time", "B04-lr10_pipe_trans", "04-anglemap_desorb",
                "04b-anglemap_desorb-sticking_source"]


@pytest.mark.parametrize("name", TRACE_MODELS)
def test_event_traces_match_oracle(gpu, name):
    m = geometry.tube_model() if name == "tube" else model_from_fixture(name)[0]
    n, k = 500, 30
    loaded(gpu, m, seed=5)
    gpu.set_trace(n, k)
    gpu.run(n, max_events=MAXEV)
    fg, kg = canonical_trace(*gpu.get_trace())
    gpu.set_trace(0, 0)
    fc, kc = canonical_trace(*CpuModel(m).trace(n, k, seed=5))
    same = np.all((fg == fc) & (kg == kc), axis=1)
    assert same.mean() >= 0.98, f"only {same.mean():.3f} of {n} particles follow the oracle's first {k} events"


# ---------------------------------------------------------------- seeded tallies
def test_same_seed_tallies_match_oracle_tube(gpu):
    m = geometry.tube_model()
    n = 20_000
    loaded(gpu, m, seed=21)
    gpu.run(n, max_events=MAXEV)
    g = gpu.read()
    c = CpuModel(m).run(n, seed=21)
    hg, hc = int(g.glob[0]["nb_mc_hit"]), int(c.glob[0]["nb_mc_hit"])
    assert abs(hg - hc) <= 1e-3 * hc
    assert g.glob[0]["nb_desorbed"] == n
    a = g.facets[0]["nb_mc_hit"].astype(np.int64)
    b = c.facets[0]["nb_mc_hit"].astype(np.int64)
    assert np.abs(a - b).sum() <= 2e-3 * b.sum()
    for fld in ("sum_v_ort", "sum_1_per_velocity"):
        assert abs(g.facets[0][fld].sum() / c.facets[0][fld].sum() - 1) < 1e-3


def test_plain_tube_transmission_and_chi2(gpu):
    # config-1 physics: GPU at 1e6 desorptions vs oracle at 2e5 (different seeds)
    m = geometry.tube_model()
    loaded(gpu, m, seed=42)
    n_g = 1_000_000
    gpu.run(n_g, max_events=MAXEV)
    g = gpu.read()
    n_c = 200_000
    c = CpuModel(m).run(n_c, seed=1234)
    z = z_two_proportions(g.facets[0]["nb_abs_equiv"][1], n_g, c.facets[0]["nb_abs_equiv"][1], n_c)
    assert abs(z) < 3
    chi2, dof, p = absorption_chi2(g.facets[0]["nb_abs_equiv"], n_g, c.facets[0]["nb_abs_equiv"], n_c)
    assert p > 1e-3


def test_beam_pipe_absorption_chi2(gpu):
    m = geometry.beam_pipe_model()
    loaded(gpu, m, seed=3)
    n_g = 200_000
    gpu.run(n_g, max_events=MAXEV)
    g = gpu.read()
    n_c = 20_000
    c = CpuModel(m).run(n_c, seed=99)
    # absorbers grouped into the 8 pumps and the 6 NEG bands: a 14-way
    # multinomial (each particle is absorbed exactly once)
    grp = np.full(m.n_facets, -1)
    grp[m.pump_ids] = np.arange(len(m.pump_ids))
    band = np.floor((m.wall_z - 20.0) / 25.0).astype(int)
    grp[m.neg_mask] = 8 + np.clip(band[m.neg_mask], 0, 5) + (m.wall_z[m.neg_mask] > 100) * 0
    zz = m.wall_z[m.neg_mask]
    starts = np.array([20.0, 45.0, 70.0, 120.0, 145.0, 170.0])
    grp[np.nonzero(m.neg_mask)[0]] = 8 + np.searchsorted(starts, zz, side="right") - 1
    ag = np.bincount(grp[grp >= 0], weights=g.facets[0]["nb_abs_equiv"][grp >= 0], minlength=14)
    ac = np.bincount(grp[grp >= 0], weights=c.facets[0]["nb_abs_equiv"][grp >= 0], minlength=14)
    chi2, dof, p = absorption_chi2(ag, n_g, ac, n_c)
    assert dof >= 12 and p > 1e-3, (chi2, dof)
    # hits per desorption: batch-means on the GPU side
    hd = g.glob[0]["nb_mc_hit"] / n_g
    hc = c.glob[0]["nb_mc_hit"] / n_c
    assert abs(hd / hc - 1) < 0.05


@pytest.mark.parametrize("name", ["01-quick_pipe_profiles_textures_2sided", "02b-timedependent_with_three_parameters",
                                  "03-anglemap_record", "05-three_structures_nonsquare_textures",
                                  "07-volume_decay", "08-wall_sojourn_time", "09-histograms", "B02-lr10_pipe_tex",
                                  "B03-lr10_pipe_prof", "B04-lr10_pipe_trans"])
def test_recordings_same_seed_match_oracle(gpu, name):
    """Textures, profiles, directions, angle maps, histograms and per-moment
    counters of one seeded run agree with the oracle's run of the same seed."""
    m, _ = model_from_fixture(name)
    n = 5000
    cap = 60   # both sides stop a particle after the same 60 events
    loaded(gpu, m, seed=77)
    gpu.run(n, max_events=cap)
    g = gpu.read()
    c = CpuModel(m).run(n, seed=77, max_events=cap)
    tot = c.facets["nb_mc_hit"].astype(np.float64).sum()
    assert np.abs(g.facets["nb_mc_hit"].astype(np.float64) - c.facets["nb_mc_hit"]).sum() <= 0.005 * tot
    pairs = [(g.textures["count_equiv"], c.textures["count_equiv"]), (g.profiles["count_equiv"], c.profiles["count_equiv"]),
             (g.directions["count"].astype(float), c.directions["count"].astype(float)),
             (g.angle_maps.astype(float), c.angle_maps.astype(float)), (g.global_hist, c.global_hist),
             (g.facet_hist, c.facet_hist)]
    for a, b in pairs:
        if b.size and np.abs(b).sum() > 0:
            assert np.abs(a - b).sum() <= 0.01 * np.abs(b).sum() + 2


COMPARE = [
    ("01-quick_pipe_profiles_textures_2sided", 4_000_000), ("02-timedependent_with_two_parameters", 1_000_000),
    ("02b-timedependent_with_three_parameters", 300_000), ("03-anglemap_record", 1_000_000),
    ("03b-anglemap_record_transparent_target", 1_000_000), ("05-three_structures_nonsquare_textures", 2_000_000),
    ("07-volume_decay", 200_000), ("08-wall_sojourn_time", 2_000_000), ("09-histograms", 4_000_000),
    ("B01-lr1000_pipe", 400_000), ("B02-lr10_pipe_tex", 4_000_000), ("B03-lr10_pipe_prof", 4_000_000),
    ("B04-lr10_pipe_trans", 2_000_000), ("04-anglemap_desorb", 1_000_000),
    ("04b-anglemap_desorb-sticking_source", 1_000_000),
]


@pytest.mark.parametrize("name,n", COMPARE)
def test_gpu_compare_with_stored_results(gpu, name, n):
    """The reference's own acceptance rule (Compare, thresholds 0.009/0.07)
    between a GPU run and the results stored in the fixture."""
    m, d = model_from_fixture(name)
    loaded(gpu, m, seed=11)
    batches = 8
    ratios = []
    for b in range(batches):
        r = gpu.run(n // batches, max_events=MAXEV)
        ratios.append(r.hits / (n // batches))
    t = gpu.read()
    r = np.array(ratios)
    se = r.std(ddof=1) / math.sqrt(batches)
    S = StoredState(d)
    cnt = {}
    g_err, f_err, fine, msgs = compare_states(TallyState(t, m), S, glob=max(0.009, 4 * se / r.mean()), loc=0.07,
                                              counts=cnt)
    assert g_err == 0 and f_err == 0, msgs[:5]
    assert fine <= 0.1 * cnt["fine_checks"] + 2, (fine, cnt)


@pytest.mark.parametrize("name", ["B02-lr10_pipe_tex", "01-quick_pipe_profiles_textures_2sided", "09-histograms",
                                  "05-three_structures_nonsquare_textures"])
def test_gpu_results_wrong_protocol(gpu, name):
    """The reference's negative control (ResultsWrong): 15 GPU runs of 300
    desorptions must differ from the stored results in at least 66 % of the
    comparisons at 0.006 / 0.05."""
    m, d = model_from_fixture(name)
    S = StoredState(d)

    def short(i):
        loaded(gpu, m, seed=500 + i)
        gpu.run(300, max_events=MAXEV)
        return TallyState(gpu.read(), m)

    det, n, comparable = results_wrong_detection(short, S)
    assert det / n >= 0.66, (det, n, comparable)


def test_transmission_vs_reference_values(gpu):
    for name, stored, sig, n in (("B02-lr10_pipe_tex", 0.190947, 3.9e-5, 10_000_000),
                                 ("01-quick_pipe_profiles_textures_2sided", 0.281945, 1.7e-5, 10_000_000),
                                 ("B01-lr1000_pipe", 5262 / 2e6, 3.6e-5, 1_000_000)):
        m, _ = model_from_fixture(name)
        loaded(gpu, m, seed=5)
        gpu.run(n, max_events=MAXEV)
        t = gpu.read()
        p = t.facets[0]["nb_abs_equiv"][1] / n
        z = (p - stored) / math.sqrt(stored * (1 - stored) / n + sig ** 2)
        assert abs(z) < 3, (name, p, stored, z)


def test_time_dependent_config3_moments(gpu):
    m = geometry.time_dependent_tube()
    loaded(gpu, m, seed=8)
    n = 200_000
    gpu.run(n, max_events=MAXEV)
    g = gpu.read()
    c = CpuModel(m).run(50_000, seed=8)
    assert g.n_slots == 101
    # per-moment wall hits per desorption agree within 4 sigma (Poisson with
    # an overdispersion allowance of 3 for correlated hits)
    gw = g.facets["nb_mc_hit"][1:, 2:3202].astype(float).sum(1) / n
    cw = c.facets["nb_mc_hit"][1:, 2:3202].astype(float).sum(1) / 50_000
    sig = np.sqrt(3 * (gw / n + cw / 50_000)) + 1e-12
    ok = cw * 50_000 > 50
    assert np.all(np.abs(gw - cw)[ok] < 4 * sig[ok])
    # profile of the transparent recording facet, all moments
    pg = g.profile_of(3202, slot=0)["count_equiv"] / n
    pc = c.profile_of(3202, slot=0)["count_equiv"] / 50_000
    assert np.corrcoef(pg, pc)[0, 1] > 0.98
    # the steady-state slot is the sum over moments plus out-of-window hits
    assert g.facets["nb_mc_hit"][0].sum() >= g.facets["nb_mc_hit"][1:].sum()


def test_angle_map_source_without_a_map_is_refused(gpu):
    from molflow_sources_amd.simulation import MolflowGpuError
    m, _ = model_from_fixture("04-anglemap_desorb")
    m.am_maps = {}
    with pytest.raises(MolflowGpuError, match="DES_ANGLEMAP"):
        gpu.load(m)


def test_angle_map_source_same_seed_tallies(gpu):
    """DES_ANGLEMAP desorption draws the same directions as the oracle: the
    same seed gives the same integer tallies."""
    m, _ = model_from_fixture("04-anglemap_desorb")
    n = 20_000
    loaded(gpu, m, seed=31)
    gpu.run(n, max_events=MAXEV)
    g = gpu.read()
    c = CpuModel(m).run(n, seed=31, max_events=MAXEV)
    diff = np.abs(g.facets["nb_mc_hit"][0].astype(np.int64) - c.facets["nb_mc_hit"][0].astype(np.int64)).sum()
    assert diff <= 1e-3 * c.glob[0]["nb_mc_hit"], diff
