"""Particles kept in flight between runs (mfg_set_carry). A particle's
events depend only on its own random stream, so splitting the work into
carry runs plus a final drain must reproduce one run that traces every
particle to its end: integer tallies exactly, float sums to rounding.
Needs an MI355X."""
import threading

import numpy as np
import pytest

from molflow_sources_amd import geometry

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000


@pytest.fixture(scope="module")
def sim():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.set_carry(False)
    s.close()


def snapshot(t):
    return {
        "hits": int(t.glob[0]["nb_mc_hit"]), "des": int(t.glob[0]["nb_desorbed"]),
        "f_hits": t.facets["nb_mc_hit"].astype(np.int64).copy(),
        "f_des": t.facets["nb_desorbed"].astype(np.int64).copy(),
        "f_abs": t.facets["nb_abs_equiv"].copy(), "f_s1v": t.facets["sum_1_per_ort_velocity"].copy(),
        "tex": t.textures["count_equiv"].copy() if t.textures.size else np.zeros(0),
        "am": t.angle_maps.copy() if t.angle_maps.size else np.zeros(0, dtype=np.uint64),
        "abs_equiv": float(t.glob[0]["nb_abs_equiv"]),
    }


def assert_same(a, b):
    assert a["hits"] == b["hits"] and a["des"] == b["des"]
    assert np.array_equal(a["f_hits"], b["f_hits"]) and np.array_equal(a["f_des"], b["f_des"])
    assert np.array_equal(a["am"], b["am"])
    assert np.allclose(a["f_abs"], b["f_abs"], rtol=1e-9, atol=1e-9)
    assert np.allclose(a["f_s1v"], b["f_s1v"], rtol=1e-9, atol=0)
    assert np.allclose(a["tex"], b["tex"], rtol=1e-9, atol=1e-12)


def one_shot(sim, n, seed):
    sim.set_carry(False)
    sim.reset(seed=seed)
    r = sim.run(n, max_events=MAXEV)
    assert r.desorbed == n and sim.in_flight() == 0
    return snapshot(sim.read())


@pytest.mark.parametrize("name", ["tube", "beam"])
def test_carry_runs_then_drain_equal_one_run(sim, name):
    if name == "tube":
        m, n, k = geometry.tube_model(), 400_000, 4
    else:
        m, n, k = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=4, n_anglemaps=2), 60_000, 3
    sim.load(m)
    sim.rebuild_accel()
    ref = one_shot(sim, n * k, seed=21)
    sim.reset(seed=21)
    sim.set_carry(True)
    total_hits, parked = 0, []
    for _ in range(k):
        r = sim.run(n, max_events=MAXEV)
        assert r.desorbed == n and r.status == 0
        total_hits += r.hits
        parked.append(sim.in_flight())
    r = sim.drain(max_events=MAXEV)
    total_hits += r.hits
    assert r.desorbed == 0 and sim.in_flight() == 0
    # the queue runs dry while lanes still hold particles
    assert max(parked) > 0
    got = snapshot(sim.read())
    assert total_hits == got["hits"]
    assert_same(got, ref)


def test_carry_across_launch_shapes_and_variants(sim):
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=2, n_anglemaps=1)
    sim.load(m)
    sim.rebuild_accel()
    ref = one_shot(sim, 200_000, seed=5)
    sim.reset(seed=5)
    sim.set_carry(True)
    # shrinking and growing the grid keeps the parked particles
    for (block, per_cu, waves) in ((64, 1, 0), (1024, 0, 0), (768, 0, 0), (256, 0, 1)):
        sim.set_launch(block, per_cu)
        sim.set_kernel_variant(waves)
        sim.run(50_000, max_events=MAXEV)
    sim.set_launch(256, 0)
    sim.set_kernel_variant(0)
    sim.drain(max_events=MAXEV)
    assert_same(snapshot(sim.read()), ref)


def test_carry_batched_variants(sim):
    vs = geometry.variant_models(4, min_facets=2000, max_facets=20000, seed=8)
    nd, seeds = [40_000, 10_000, 25_000, 5_000], [31, 32, 33, 34]
    sim.load_variants(vs)
    sim.rebuild_accel()
    sim.set_carry(False)
    sim.reset()
    sim.run_variants([3 * x for x in nd], seeds, max_events=MAXEV)
    ref = [snapshot(sim.read(variant=v)) for v in range(4)]
    sim.reset()
    sim.set_carry(True)
    for _ in range(3):
        sim.run_variants(nd, seeds, max_events=MAXEV)
    sim.drain(max_events=MAXEV)
    for v in range(4):
        assert_same(snapshot(sim.read(variant=v)), ref[v])


def test_carry_graph_replays(sim):
    m = geometry.tube_model()
    sim.load(m)
    sim.rebuild_accel()
    ref = one_shot(sim, 5 * 200_000, seed=13)
    sim.reset(seed=13)
    sim.set_carry(True)
    sim.enable_graph(True)
    for _ in range(5):
        sim.run(200_000, max_events=MAXEV)
    sim.enable_graph(False)
    sim.drain(max_events=MAXEV)
    assert_same(snapshot(sim.read()), ref)


def test_carry_stop_parks_particles(sim):
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=2)
    sim.set_carry(True)
    t = threading.Timer(0.3, sim.request_stop)
    t.start()
    r = sim.run(50_000_000, max_events=MAXEV)
    t.join()
    assert r.stopped == 1 and 0 < r.desorbed < 50_000_000
    assert sim.in_flight() > 0
    sim.drain(max_events=MAXEV)
    g = sim.read().glob[0]
    # every started particle ended after the drain (no leaks in this model)
    assert g["nb_desorbed"] == r.desorbed
    assert g["nb_abs_equiv"] == r.desorbed


def test_time_limit_ends_runs_whose_particles_never_end(sim):
    # nothing absorbs and no event cap: only the launch time limit, forced
    # short from the host, ends these runs
    m = geometry.tube_model()
    m.facets["sticking"][:2] = 0.0
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=4)
    sim.set_time_limit(0.25)
    try:
        sim.set_carry(False)
        r1 = sim.run(50_000_000, max_events=0)
        assert r1.stopped == 2 and 0 < r1.desorbed < 50_000_000 and r1.hits > 0
        assert sim.in_flight() == 0
        sim.set_carry(True)
        r2 = sim.run(50_000_000, max_events=0)
        # every particle started is still alive, and parked
        assert r2.stopped == 2 and 0 < r2.desorbed < 50_000_000
        assert sim.in_flight() == r2.desorbed
        r3 = sim.run(0, max_events=0)
        assert r3.stopped == 2 and r3.desorbed == 0 and r3.hits > 0
        assert sim.in_flight() == r2.desorbed
        g = sim.read().glob[0]
        assert g["nb_desorbed"] == r1.desorbed + r2.desorbed
        assert g["nb_mc_hit"] == r1.hits + r2.hits + r3.hits
        assert g["nb_abs_equiv"] == 0.0
    finally:
        sim.set_time_limit(60.0)
        sim.set_carry(False)
    # the default limit leaves an ordinary run alone
    sim.load(geometry.tube_model())
    sim.rebuild_accel()
    sim.reset(seed=4)
    r = sim.run(100_000, max_events=MAXEV)
    assert r.stopped == 0 and r.desorbed == 100_000


def test_carry_refused_with_adjoint(sim):
    from molflow_sources_amd.simulation import MolflowGpuError
    m = geometry.beam_pipe_model(n_az=24, n_ax=60, port_rings=3, n_textures=0, n_anglemaps=0)
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-3
    grp, base, tg = geometry.adjoint_setup(m, n_groups=4, n_targets=8)
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=1)
    sim.set_carry(False)
    sim.adjoint_setup(grp, tg)
    with pytest.raises(MolflowGpuError):
        sim.set_carry(True)


@pytest.mark.parametrize("name", ["tube", "beam"])
def test_take_after_each_carry_run_sums_to_one_run(sim, name):
    """The periodic-report path of the boundary: carry on, k runs with a
    take (add into the caller's state, then clear the device tallies)
    after each, a drain and a last take. Parked particles survive every
    take, so the summed state equals one uninterrupted run exactly in its
    integer tallies (the UpdateHitsAndLog contract, Particle.cpp:1294-1312)."""
    if name == "tube":
        m, n, k = geometry.tube_model(), 300_000, 4
    else:
        m, n, k = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=4, n_anglemaps=2), 50_000, 3
    sim.load(m)
    sim.rebuild_accel()
    ref = one_shot(sim, n * k, seed=17)
    sim.reset(seed=17)
    sim.set_carry(True)
    acc = None
    parked = []
    for _ in range(k):
        r = sim.run(n, max_events=MAXEV)
        assert r.desorbed == n
        parked.append(sim.in_flight())
        acc = sim.take(acc)
        # the device state is empty after a take, the parked particles are not
        assert int(sim.read().glob[0]["nb_mc_hit"]) == 0
        assert sim.in_flight() == parked[-1]
    assert max(parked) > 0
    sim.drain(max_events=MAXEV)
    acc = sim.take(acc)
    assert sim.in_flight() == 0
    assert_same(snapshot(acc), ref)


def test_time_limit_keeps_parked_particles_it_never_resumed(sim):
    """Particles parked by a large grid, resumed by a much smaller one that
    hits its time limit before it claims them all: the ones it never
    claimed stay parked (none is lost), and in-flight plus ended equals
    what was parked before."""
    m = geometry.tube_model()
    m.facets["sticking"][:2] = 0.0
    m.facets["sticking"][2:] = 2e-4        # particles end, but only after thousands of bounces
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=9)
    sim.set_carry(True)
    try:
        sim.set_time_limit(0.2)
        r1 = sim.run(20_000_000, max_events=0)
        assert r1.stopped == 2
        before = sim.in_flight()
        g0 = sim.read().glob[0]
        ended0 = int(g0["nb_abs_equiv"]) + int(g0["nb_leak_total"])
        assert before + ended0 == r1.desorbed
        sim.set_launch(64, 1)                  # 64 lanes per CU: far fewer than parked
        sim.set_time_limit(0.05)
        r2 = sim.run(0, max_events=0)
        assert r2.stopped == 2 and r2.desorbed == 0
        after = sim.in_flight()
        g1 = sim.read().glob[0]
        ended1 = int(g1["nb_abs_equiv"]) + int(g1["nb_leak_total"])
        assert after > 64 * 256, "the small grid should not have claimed every parked particle"
        assert after + (ended1 - ended0) == before
        # a full drain ends every one of them
        sim.set_launch(256, 0)
        sim.set_time_limit(60.0)
        sim.drain(max_events=0)
        g2 = sim.read().glob[0]
        assert int(g2["nb_abs_equiv"]) + int(g2["nb_leak_total"]) == r1.desorbed
    finally:
        sim.set_launch(256, 0)
        sim.set_time_limit(60.0)
        sim.set_carry(False)
