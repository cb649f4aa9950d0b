"""The five workloads at their configured size, on the GPU through the C ABI,
against the oracle. Needs an MI355X.

  1 tube, 3,200 wall quads: per-facet wall hits, batch means, chi-square
    over every wall (GPU and oracle on different seeds);
  2 beam pipe, ~100k triangles: its 16 textures and 4 angle maps, same
    seed with a per-particle event cap;
  3 time-dependent tube, 100 moments: per-moment facet counters and the
    recording facet's per-moment profiles, same seed, slot by slot;
  4 256 uneven beam-pipe variants (2k-150k facets) in one launch: four of
    them against the oracle, same seed;
  5 full beam pipe with 32 sticking groups x 64 targets: path scores and
    the Jacobian against the oracle's, same seed with an event cap.
"""
import math

import numpy as np
import pytest

from molflow_sources_amd import geometry
from oracle.mfcpu import CpuModel

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000
THREADS = 16


@pytest.fixture(scope="module")
def sim():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.close()


def loaded(sim, m, seed):
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=seed)


def test_config1_wall_hits_chi2_over_every_wall(sim):
    """Hits per desorption on each of the 3,200 walls: batch means on both
    sides (GPU 16 x 125k, oracle 16 x 12.5k desorptions, different seeds),
    z^2 of the difference per wall. Each z^2 follows a squared Student t of
    about 30 degrees of freedom (mean 1.07); a wrong reflection, source or
    intersection law moves the walls' hit profile and the mean far up."""
    m = geometry.tube_model()
    walls = slice(2, 2 + 3200)
    nb = 16

    def gpu_run(b, n):
        sim.reset(seed=1000 + b)
        sim.run(n, max_events=MAXEV)
        return sim.read()

    def batches(run, n):
        return np.array([run(b, n).facets[0]["nb_mc_hit"][walls].astype(np.float64) / n for b in range(nb)])

    sim.load(m)
    sim.rebuild_accel()
    cm = CpuModel(m)
    G = batches(gpu_run, 125_000)
    C = batches(lambda b, n: cm.run(n, seed=2000 + b, threads=THREADS), 12_500)
    z2 = (G.mean(0) - C.mean(0)) ** 2 / (G.var(0, ddof=1) / nb + C.var(0, ddof=1) / nb)
    assert z2.size == 3200
    assert 0.85 < z2.mean() < 1.35, z2.mean()
    assert z2.max() < 45, (int(np.argmax(z2)), float(z2.max()))


def test_config2_textures_and_angle_maps_same_seed(sim):
    """Full beam pipe with its 16 textures (256^2) and 4 angle maps
    (90 x 180): one seeded run of 20k particles capped at 200 events each,
    GPU and oracle follow the same paths, so every texture and map agrees
    cell by cell up to the few paths that part on rounding."""
    m = geometry.beam_pipe_model()
    n, cap = 20_000, 200
    loaded(sim, m, seed=91)
    sim.run(n, max_events=cap)
    g = sim.read()
    cm = CpuModel(m)
    c = cm.run(n, seed=91, max_events=cap, threads=THREADS)
    c.offsets = cm.offsets()
    off = g.offsets
    for k in off:
        assert np.array_equal(off[k], c.offsets[k]), k
    assert g.layout.n_tex_cells == 16 * 256 * 256 and g.layout.n_angle_cells == 4 * 90 * 180
    tot = c.facets["nb_mc_hit"].astype(np.float64).sum()
    assert np.abs(g.facets["nb_mc_hit"].astype(np.float64) - c.facets["nb_mc_hit"]).sum() <= 0.005 * tot
    tex_facets = np.nonzero(off["texture"] >= 0)[0]
    am_facets = np.nonzero(off["anglemap"] >= 0)[0]
    assert len(tex_facets) == 16 and len(am_facets) == 4
    for f in tex_facets:
        a = g.texture_of(f, 256, 256)["count_equiv"]
        b = c.texture_of(f, 256, 256)["count_equiv"]
        assert b.sum() > 0
        assert np.abs(a - b).sum() <= 0.01 * b.sum() + 2, f
    for f in am_facets:
        o = int(off["anglemap"][f])
        a = g.angle_maps[o:o + 90 * 180].astype(np.float64)
        b = c.angle_maps[o:o + 90 * 180].astype(np.float64)
        assert b.sum() > 0
        assert np.abs(a - b).sum() <= 0.01 * b.sum() + 2, f


def test_config3_moments_slot_by_slot_same_seed(sim):
    """Time-dependent tube (100 moments): same seed, every particle run to
    its end; each moment slot's per-facet hit and absorption counts and the
    recording facet's profile in each slot agree with the oracle's."""
    m = geometry.time_dependent_tube()
    n = 20_000
    loaded(sim, m, seed=8)
    sim.run(n, max_events=MAXEV)
    g = sim.read()
    c = CpuModel(m).run(n, seed=8, threads=THREADS)
    assert g.n_slots == c.n_slots == 101
    rec = 3202
    for s in range(g.n_slots):
        a = g.facets[s]["nb_mc_hit"].astype(np.int64)
        b = c.facets[s]["nb_mc_hit"].astype(np.int64)
        assert np.abs(a - b).sum() <= 0.005 * b.sum() + 4, s
        a = g.facets[s]["nb_abs_equiv"]
        b = c.facets[s]["nb_abs_equiv"]
        assert np.abs(a - b).sum() <= 0.005 * b.sum() + 4, s
        pa = g.profile_of(rec, slot=s)["count_equiv"]
        pb = c.profile_of(rec, slot=s)["count_equiv"]
        assert np.abs(pa - pb).sum() <= 0.01 * pb.sum() + 4, s


def test_config4_256_variants_against_the_oracle(sim):
    """256 uneven variants (2k-150k facets, 7.8M in all) in one launch;
    the smallest, the largest and two others run 20k desorptions each (the
    rest 300) and their integer tallies match the oracle's run of the same
    seed. Every particle stops after 100 events on both sides: the low-
    sticking variants make paths of thousands of hits, along which rounding
    differences of the two sides' libm (the bounce's sine and cosine)
    eventually send the same particle to different facets."""
    vs = geometry.variant_models(256, seed=42)
    sizes = np.array([v.n_facets for v in vs])
    assert len(vs) == 256 and sizes.min() < 3000 and sizes.max() > 100_000
    check = [int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 200]
    nd = np.full(256, 300, dtype=np.int64)
    nd[check] = 20_000
    seeds = 5000 + np.arange(256)
    sim.load_variants(vs)
    sim.rebuild_accel()
    sim.reset()
    cap = 100
    r = sim.run_variants(nd, seeds, max_events=cap)
    assert r.desorbed == nd.sum()
    for v in check:
        tb = sim.read(variant=v)
        tc = CpuModel(vs[v]).run(int(nd[v]), seed=int(seeds[v]), threads=THREADS, max_events=cap)
        assert tb.glob[0]["nb_desorbed"] == tc.glob[0]["nb_desorbed"] == nd[v]
        hg, hc = int(tb.glob[0]["nb_mc_hit"]), int(tc.glob[0]["nb_mc_hit"])
        assert abs(hg - hc) <= 1e-3 * hc, (v, hg, hc)
        a = tb.facets[0]["nb_mc_hit"].astype(np.int64)
        b = tc.facets[0]["nb_mc_hit"].astype(np.int64)
        assert np.abs(a - b).sum() <= 2e-3 * b.sum(), v
        assert np.array_equal(tb.facets[0]["nb_desorbed"], tc.facets[0]["nb_desorbed"]), v
        a = tb.facets[0]["nb_abs_equiv"]
        b = tc.facets[0]["nb_abs_equiv"]
        assert np.abs(a - b).sum() <= 4e-3 * b.sum() + 2, v


def config5_model():
    m = geometry.beam_pipe_model()
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-6
    grp, base, tg = geometry.adjoint_setup(m, n_groups=32, n_targets=64, seed=42)
    for g, s in enumerate(base):
        m.facets["sticking"][grp == g] = s
    return m, grp, base, tg


def test_config5_path_scores_match_the_oracle(sim):
    """Full beam pipe, 32 groups x 64 targets, low-flux mode: the first 40
    events of 4,000 particles of the same seed give the same J0 and
    Jacobian on the GPU and in the oracle."""
    m, grp, base, tg = config5_model()
    n = 4000
    sim.load(m)
    sim.rebuild_accel()
    sim.reset(seed=13)
    sim.adjoint_setup(grp, tg)
    sim.adjoint_set_sticking(base)
    sim.run(n, max_events=40)
    j0, jac = sim.adjoint_read()
    cj0, cjac = CpuModel(m).adjoint(n, grp, base, tg, seed=13, max_events=40)
    assert jac.shape == (64, 32)
    assert np.abs(cj0).sum() > 0 and np.abs(cjac).sum() > 0
    assert np.abs(j0 - cj0).sum() <= 0.01 * np.abs(cj0).sum()
    assert np.abs(jac - cjac).sum() <= 0.01 * np.abs(cjac).sum()


def test_config5_gradient_matches_paired_differences(sim):
    """Full beam-pipe mesh, 32 groups x 64 targets, cutoff 1e-6: every
    d(up . J0)/ds_g of the path-score Jacobian (with the low-flux cutoff's
    term) against Richardson-extrapolated central differences (h and 2h,
    h = 1e-2 s_g) of the same particles, over 12 independent particle
    blocks. The 128 perturbed models run as the variants of one batched
    launch per block (same seed and particle numbers as the gradient's run,
    so the same paths; J0 is the targets' hit-equivalents). Each group's
    paired z (mean difference over its own standard error) follows
    Student's t with 11 degrees of freedom when the gradient is unbiased:
    all 32 lie within 4.5 (3 % chance otherwise) and their mean square is
    below 2.4 (mean 11/9, 3 sigma above). Requiring |z| <= 2 for all 32
    would fail about nine times in ten without any bias (0.93^32 = 0.1).
    Textures and angle maps are left off: they record, and do not move,
    particles."""
    import dataclasses
    from molflow_sources_amd.simulation import GpuSimulation
    m = geometry.beam_pipe_model(n_textures=0, n_anglemaps=0)
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-6
    grp, base, tg = geometry.adjoint_setup(m, n_groups=32, n_targets=64, seed=42)
    for g, s in enumerate(base):
        m.facets["sticking"][grp == g] = s
    per, nb = 4_000, 12
    up = np.random.default_rng(42).normal(size=len(tg))
    sim.load(m)
    sim.rebuild_accel()
    sim.adjoint_setup(grp, tg)
    adj = []
    for b in range(nb):
        sim.adjoint_set_sticking(base)
        sim.reset(seed=42, particle_offset=b * per)
        sim.run(per, max_events=MAXEV)
        adj.append((up @ sim.adjoint_read()[1]) / per)
    adj = np.array(adj)
    steps = (-2, -1, 1, 2)
    variants = []
    for g in range(len(base)):
        for q in steps:
            F = m.facets.copy()
            F["sticking"][grp == g] = base[g] + q * 1e-2 * base[g]
            variants.append(dataclasses.replace(m, facets=F))
    fds = GpuSimulation(0)
    try:
        fds.load_variants(variants)
        fds.rebuild_accel()
        fd = np.zeros_like(adj)
        nv = len(variants)
        for b in range(nb):
            fds.reset(seed=42, particle_offset=b * per)
            fds.run_variants(np.full(nv, per), np.full(nv, 42), max_events=MAXEV)
            J = np.array([up @ fds.read(variant=v).facets[0]["nb_hit_equiv"][tg] / per for v in range(nv)])
            J = J.reshape(len(base), len(steps))
            h = 1e-2 * base
            d1 = (J[:, 2] - J[:, 1]) / (2 * h)
            d2 = (J[:, 3] - J[:, 0]) / (4 * h)
            fd[b] = (4 * d1 - d2) / 3
            print(f"paired differences: block {b + 1}/{nb} done", flush=True)
    finally:
        fds.close()
    d = fd - adj
    z = d.mean(0) / (d.std(0, ddof=1) / math.sqrt(nb))
    print("paired z:", np.round(z, 2).tolist(), flush=True)
    assert np.all(np.abs(z) < 4.5), np.round(z, 2)
    assert np.mean(z ** 2) < 2.4, np.round(z, 2)
