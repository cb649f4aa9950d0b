"""Batched variants, launch-shape independence, hipGraph replays, stop
requests and the sticking adjoint. Needs an MI355X."""
import math
import threading
import time

import numpy as np
import pytest

from molflow_sources_amd import geometry
from oracle.mfcpu import CpuModel

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000


@pytest.fixture(scope="module")
def sim():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.close()


def int_tallies(t):
    return (int(t.glob[0]["nb_mc_hit"]), int(t.glob[0]["nb_desorbed"]), t.facets["nb_mc_hit"].astype(np.int64).copy(),
            t.facets["nb_desorbed"].astype(np.int64).copy())


def test_batched_variants_equal_plain_runs(sim):
    vs = geometry.variant_models(6, min_facets=2000, max_facets=30000, seed=4)
    nd = [3000, 500, 2000, 800, 1500, 1000]
    seeds = [101, 102, 103, 104, 105, 106]
    sim.load_variants(vs)
    sim.rebuild_accel()
    sim.reset()
    r = sim.run_variants(nd, seeds, max_events=MAXEV)
    assert r.desorbed == sum(nd)
    from molflow_sources_amd.simulation import GpuSimulation
    p = GpuSimulation(0)
    for v, m in enumerate(vs):
        tb = sim.read(variant=v)
        p.load(m)
        p.rebuild_accel()
        p.reset(seed=seeds[v])
        p.run(nd[v], max_events=MAXEV)
        tp = p.read()
        a, b = int_tallies(tb), int_tallies(tp)
        # identical RNG streams: the batched launch reproduces the plain run
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    p.close()


def test_launch_shape_does_not_change_results(sim):
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=4, n_anglemaps=2)
    sim.load(m)
    sim.rebuild_accel()
    res = []
    for block, per_cu in ((256, 0), (64, 3), (128, 1), (768, 0), (1024, 0)):
        sim.set_launch(block, per_cu)
        sim.reset(seed=9)
        sim.run(4000, max_events=MAXEV)
        t = sim.read()
        res.append(int_tallies(t) + (t.angle_maps.copy(), t.textures["count_equiv"].copy()))
    sim.set_launch(256, 0)
    for other in res[1:]:
        assert res[0][0] == other[0] and np.array_equal(res[0][2], other[2]) and np.array_equal(res[0][4], other[4])
        assert np.allclose(res[0][5], other[5], rtol=1e-9, atol=1e-9)
    # block sizes other than multiples of 64 up to 256, 768 and 1024 are refused
    from molflow_sources_amd.simulation import MolflowGpuError
    for bad in (512, 320, 96 + 1, 2048):
        with pytest.raises(MolflowGpuError):
            sim.set_launch(bad, 0)


def test_runs_continue_particle_streams(sim):
    m = geometry.tube_model()
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=3)
    sim.run(5000, max_events=MAXEV)
    sim.run(5000, max_events=MAXEV)
    two = int_tallies(sim.read())
    sim.reset(seed=3)
    sim.run(10000, max_events=MAXEV)
    one = int_tallies(sim.read())
    assert two[0] == one[0] and np.array_equal(two[2], one[2])
    # resume offset: particles 5000.. of a fresh context equal the second half
    sim.reset(seed=3, particle_offset=5000)
    sim.run(5000, max_events=MAXEV)
    half = int_tallies(sim.read())
    sim.reset(seed=3)
    sim.run(5000, max_events=MAXEV)
    first = int_tallies(sim.read())
    assert half[0] + first[0] == one[0]


def test_graph_replays_match_eager(sim):
    vs = geometry.variant_models(4, min_facets=2000, max_facets=8000, seed=6)
    nd, seeds = [1000, 2000, 500, 1500], [1, 2, 3, 4]
    sim.load_variants(vs)
    sim.rebuild_accel()
    sim.reset()
    for _ in range(3):
        sim.run_variants(nd, seeds, max_events=MAXEV)
    eager = [int_tallies(sim.read(variant=v)) for v in range(4)]
    sim.enable_graph(True)
    sim.reset()
    for _ in range(3):
        sim.run_variants(nd, seeds, max_events=MAXEV)
        time.sleep(0.01)
    graph = [int_tallies(sim.read(variant=v)) for v in range(4)]
    sim.enable_graph(False)
    for a, b in zip(eager, graph):
        assert a[0] == b[0] and np.array_equal(a[2], b[2])


def test_request_stop_ends_run_early(sim):
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=1)
    t = threading.Timer(0.3, sim.request_stop)
    t.start()
    r = sim.run(50_000_000, max_events=MAXEV)
    t.join()
    assert r.stopped == 1 and 0 < r.desorbed < 50_000_000
    g = sim.read().glob[0]
    # accounting: every started particle ended (absorbed, leaked or dropped)
    assert g["nb_desorbed"] == r.desorbed
    assert abs(g["nb_abs_equiv"] - r.desorbed) <= 0


def adjoint_model():
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-6
    g, s, t = geometry.adjoint_setup(m, n_groups=8, n_targets=16, seed=5)
    return m, g, s, t


def test_adjoint_matches_oracle_path_scores(sim):
    """Same seed, first 40 events of every particle: GPU and oracle follow
    the same paths, so J0 and the path-score Jacobian agree closely."""
    m, grp, base, tg = adjoint_model()
    n = 4000
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=13)
    sim.adjoint_setup(grp, tg)
    sim.adjoint_set_sticking(base)
    sim.run(n, max_events=40)
    j0, jac = sim.adjoint_read()
    cm = CpuModel(m)
    cj0, cjac = cm.adjoint(n, grp, base, tg, seed=13, max_events=40)
    assert np.abs(j0 - cj0).sum() <= 0.01 * np.abs(cj0).sum()
    assert np.abs(jac - cjac).sum() <= 0.01 * np.abs(cjac).sum()
    up = np.random.default_rng(1).normal(size=len(tg))
    g = sim.adjoint_backward(up)
    assert np.allclose(g, up @ jac / n, rtol=1e-9, atol=1e-12)


def test_adjoint_one_block_per_cu_matches_256_lane_blocks(sim):
    """The low-flux adjoint kernel in one 768-lane block per CU follows the
    same particles as 256-lane blocks: the same J0 and Jacobian up to the
    order of the FP64 sums."""
    m, grp, base, tg = adjoint_model()
    sim.load(m)
    sim.rebuild_accel()
    sim.adjoint_setup(grp, tg)
    out = []
    for block in (256, 768):
        sim.set_launch(block, 0)
        sim.reset(seed=21)
        sim.adjoint_set_sticking(base)
        sim.run(6000, max_events=MAXEV)
        out.append(sim.adjoint_read())
    sim.set_launch(256, 0)
    (a0, aj), (b0, bj) = out
    assert np.allclose(a0, b0, rtol=1e-9, atol=1e-12)
    assert np.allclose(aj, bj, rtol=1e-9, atol=1e-12)


def test_adjoint_full_paths_statistically(sim):
    """Full low-flux paths: GPU and oracle agree within 4 sigma (batch means)."""
    m, grp, base, tg = adjoint_model()
    sim.load(m)
    sim.rebuild_accel()
    sim.adjoint_setup(grp, tg)
    sim.adjoint_set_sticking(base)
    up = np.random.default_rng(3).normal(size=len(tg))

    def gpu_batches(k, n):
        out = []
        for b in range(k):
            sim.reset(seed=100 + b)
            sim.adjoint_set_sticking(base)
            sim.run(n, max_events=MAXEV)
            j0, jac = sim.adjoint_read()
            out.append(np.concatenate([j0 / n, up @ jac / n]))
        return np.array(out)

    cm = CpuModel(m)

    def cpu_batches(k, n):
        out = []
        for b in range(k):
            j0, jac = cm.adjoint(n, grp, base, tg, seed=500 + b)
            out.append(np.concatenate([j0 / n, up @ jac / n]))
        return np.array(out)

    G = gpu_batches(8, 20000)
    Cc = cpu_batches(8, 1000)
    mg, sg = G.mean(0), G.std(0, ddof=1) / math.sqrt(8)
    mc, sc = Cc.mean(0), Cc.std(0, ddof=1) / math.sqrt(8)
    z = (mg - mc) / np.sqrt(sg ** 2 + sc ** 2 + 1e-300)
    assert np.all(np.abs(z) < 4.5), z


def test_adjoint_matches_central_differences(sim):
    """d J0 / d s_g from the path-score Jacobian (with the low-flux cutoff's
    term) vs Richardson-extrapolated central differences (h and 2h,
    h = 1e-2 s_g) with common random numbers, over 8 independent particle
    blocks, for the first, a middle and the last group: each paired z is
    within 3.5."""
    m, grp, base, tg = adjoint_model()
    per, nb = 4_000, 8
    groups = (0, len(base) // 2, len(base) - 1)
    sim.load(m)
    sim.rebuild_accel()
    sim.adjoint_setup(grp, tg)
    up = np.random.default_rng(2).normal(size=len(tg))

    def forward(s, b):
        sim.adjoint_set_sticking(s)
        sim.reset(seed=17, particle_offset=b * per)
        sim.run(per, max_events=MAXEV)
        j0, jac = sim.adjoint_read()
        return float(up @ j0) / per, (up @ jac) / per

    adj = np.array([forward(base, b)[1] for b in range(nb)])[:, groups]
    fd = np.zeros_like(adj)
    for j, gi in enumerate(groups):
        h = 1e-2 * base[gi]
        for b in range(nb):
            f = {}
            for q in (-2, -1, 1, 2):
                sv = base.copy()
                sv[gi] += q * h
                f[q] = forward(sv, b)[0]
            d1 = (f[1] - f[-1]) / (2 * h)
            d2 = (f[2] - f[-2]) / (4 * h)
            fd[b, j] = (4 * d1 - d2) / 3
        print(f"paired differences: group {gi} done", flush=True)
    d = fd - adj
    z = d.mean(0) / (d.std(0, ddof=1) / math.sqrt(nb))
    assert np.all(np.abs(z) < 3.5), z


def test_bvh_width_and_autotune_do_not_change_results(sim):
    """The traversal order (BVH width 4 or 8, leaves of 1 or 2 facets), the kernel variant, the launch
    shape and the hot-facet privatisation chosen by the autotuner change no
    particle's events: integer tallies are identical for one seed."""
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=4, n_anglemaps=2)
    sim.load(m)
    res = []
    for width, leaf in ((8, 2), (4, 2), (8, 1)):
        sim.rebuild_accel(width, leaf)
        sim.reset(seed=12)
        sim.run(20_000, max_events=MAXEV)
        t = sim.read()
        res.append(int_tallies(t) + (t.angle_maps.copy(),))
    t = sim.autotune(20_000)
    assert t.bvh_width in (4, 8) and t.leaf_size in (1, 2) and t.hits_per_s > 0
    sim.reset(seed=12)
    sim.run(20_000, max_events=MAXEV)
    tt = sim.read()
    res.append(int_tallies(tt) + (tt.angle_maps.copy(),))
    sim.set_kernel_variant(0)
    sim.set_launch(256, 0)
    for other in res[1:]:
        assert res[0][0] == other[0] and res[0][1] == other[1]
        assert np.array_equal(res[0][2], other[2]) and np.array_equal(res[0][3], other[3])
        assert np.array_equal(res[0][4], other[4])


def test_hot_facets_from_autotune_keep_tube_tallies(sim):
    m = geometry.tube_model()
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=4)
    sim.run(200_000, max_events=MAXEV)
    a = sim.read()
    t = sim.autotune(100_000)
    assert t.n_hot >= 2          # inlet and outlet take every desorption and absorption
    sim.reset(seed=4)
    sim.run(200_000, max_events=MAXEV)
    b = sim.read()
    sim.set_kernel_variant(0)
    sim.set_launch(256, 0)
    assert np.array_equal(a.facets["nb_mc_hit"], b.facets["nb_mc_hit"])
    assert np.array_equal(a.facets["nb_desorbed"], b.facets["nb_desorbed"])
    for f in ("nb_abs_equiv", "sum_1_per_ort_velocity", "sum_v_ort", "sum_1_per_velocity"):
        assert np.allclose(a.facets[f], b.facets[f], rtol=1e-12, atol=0)


def test_adjoint_setup_checks_group_sticking(sim):
    """A group whose facets disagree on sticking is refused; a group whose
    shared sticking is 0 or 1 (where -n/(1-s) or 1/s is undefined) blocks
    runs until mfg_adjoint_set_sticking gives it a value inside (0,1)."""
    from molflow_sources_amd.simulation import MolflowGpuError
    m, grp, base, tg = adjoint_model()
    g0 = np.nonzero(grp == 0)[0]
    m2 = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    m2.settings.low_flux_mode = 1
    m2.facets["sticking"][g0[0]] = 0.02    # the rest of group 0 keeps 0.01
    sim.load(m2)
    sim.rebuild_accel()
    with pytest.raises(MolflowGpuError, match="mixes facets"):
        sim.adjoint_setup(grp, tg)
    for s0 in (0.0, 1.0):
        m3 = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
        m3.settings.low_flux_mode = 1
        m3.facets["sticking"][grp == 1] = s0
        sim.load(m3)
        sim.rebuild_accel()
        sim.reset(seed=3)
        sim.adjoint_setup(grp, tg)
        with pytest.raises(MolflowGpuError, match="inside"):
            sim.run(100, max_events=100)
        sim.adjoint_set_sticking(base)
        sim.run(2000, max_events=200)
        j0, jac = sim.adjoint_read()
        assert np.all(np.isfinite(jac)) and np.all(np.isfinite(j0))
