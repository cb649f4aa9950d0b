"""The HIP tracer on the models of tests/feature_models.py: the branches the
reference's fixtures leave open (specular / cos^N / mixed reflection,
partial accommodation, moving facets, forces, scattering, outgassing maps,
angular and ortho/tangential velocity profiles, teleports). The oracle's
behaviour on these models is pinned by tests/test_oracle_features.py; here
the GPU must follow the oracle particle by particle on the same seed, and
reproduce the exact identities on its own. Needs an MI355X."""
import math

import numpy as np
import pytest

import feature_models as FM
from helpers import canonical_trace
from molflow_sources_amd import _abi as A
from oracle.mfcpu import CpuModel

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000


@pytest.fixture(scope="module")
def gpu():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.close()


def _nm(**kw):
    return FM.with_settings(use_maxwell=0, gas_mass=28.0, **kw)


FEATURES = {
    "diffuse": lambda: FM.thin_box(floor=dict(), log_limit=20_000),
    "specular": lambda: FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=1.0),
                                    source=dict(desorb_type=A.DES_UNIFORM), log_limit=20_000),
    "cos_n": lambda: FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=0.0, cosine_exponent=3.0),
                                 log_limit=20_000),
    "mixed": lambda: FM.thin_box(floor=dict(diffuse_part=0.3, specular_part=0.3, cosine_exponent=3.0),
                                 source=dict(desorb_type=A.DES_UNIFORM), log_limit=20_000),
    "accommodation": lambda: FM.thin_box(floor=dict(accomodation=0.5, temperature=600.0),
                                         source=dict(temperature=300.0), settings=_nm(), log_limit=20_000),
    "accommodation_maxwell": lambda: FM.thin_box(floor=dict(accomodation=0.3, temperature=900.0),
                                                 source=dict(temperature=300.0), log_limit=20_000),
    "translation": lambda: FM.thin_box(floor=dict(is_moving=1),
                                       settings=_nm(motion_type=1, motion_vector2=(25.0, -10.0, 0.0)),
                                       log_limit=20_000),
    "rotation": lambda: FM.thin_box(floor=dict(is_moving=1),
                                    settings=_nm(motion_type=2, motion_vector1=(30.0, -20.0, 0.0),
                                                 motion_vector2=(0.0, 0.0, 400.0)), log_limit=20_000),
    "forces": lambda: FM.force_box(),
    "scatter": lambda: FM.scatter_beam(mfp=2.0, mass_ratio=1.0),
    "scatter_cutoff": lambda: FM.scatter_beam(mfp=0.5, mass_ratio=0.7, cutoff=60.0),
    "og_map": lambda: FM.og_map_plate(ratio=1.3)[0],
    "angular_profile": lambda: FM.thin_box(floor=dict(), ceiling=dict(profile_type=A.PROFILE_ANGULAR)),
    "ortho_profile": lambda: FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=0.0, cosine_exponent=3.0,
                                                    accomodation=0.99, temperature=10_000.0),
                                         ceiling=dict(profile_type=A.PROFILE_ORT_VELOCITY, temperature=1324.0),
                                         settings=_nm()),
    "tangential_profile": lambda: FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=0.0, cosine_exponent=3.0,
                                                         accomodation=0.99, temperature=10_000.0),
                                              ceiling=dict(profile_type=A.PROFILE_TAN_VELOCITY, temperature=1324.0),
                                              settings=_nm()),
    "teleport_tubes": lambda: FM.teleport_tubes(rotate=True),
    "teleport_back": lambda: FM.teleport_back_chain(),
}


def loaded(gpu, m, seed):
    gpu.load(m)
    gpu.rebuild_accel()
    gpu.reset(seed=seed)


@pytest.mark.parametrize("name", sorted(FEATURES))
def test_feature_event_traces_match_oracle(gpu, name):
    """Same seed, first 40 events of 600 particles: the event kinds and
    facets of the GPU run equal the oracle's for (nearly) every particle."""
    m = FEATURES[name]()
    n, k = 600, 40
    loaded(gpu, m, seed=5)
    gpu.set_trace(n, k)
    gpu.run(n, max_events=MAXEV)
    fg, kg = canonical_trace(*gpu.get_trace())
    gpu.set_trace(0, 0)
    fc, kc = canonical_trace(*CpuModel(m).trace(n, k, seed=5))
    same = np.all((fg == fc) & (kg == kc), axis=1)
    assert same.mean() >= 0.98, f"{name}: only {same.mean():.3f} of {n} particles follow the oracle"
    # the models exercise what they are named for
    want = {"scatter": A.EV_SCATTER, "scatter_cutoff": A.EV_SCATTER, "teleport_tubes": A.EV_TELEPORT,
            "teleport_back": A.EV_TELEPORT}.get(name)
    if want is not None:
        assert (kc == want).any()


@pytest.mark.parametrize("name", sorted(FEATURES))
def test_feature_tallies_same_seed_match_oracle(gpu, name):
    """Every tally array of one seeded run (facet counters, forces,
    textures, profiles, logs) agrees with the oracle's run of the same seed."""
    m = FEATURES[name]()
    n = 20_000
    loaded(gpu, m, seed=77)
    gpu.run(n, max_events=MAXEV)
    g = gpu.read()
    c = CpuModel(m).run(n, seed=77, max_events=MAXEV)
    assert int(g.glob[0]["nb_desorbed"]) == n
    hc = int(c.glob[0]["nb_mc_hit"])
    assert abs(int(g.glob[0]["nb_mc_hit"]) - hc) <= 2e-3 * hc + 2
    tot = c.facets["nb_mc_hit"].astype(np.float64).sum()
    assert np.abs(g.facets["nb_mc_hit"].astype(np.float64) - c.facets["nb_mc_hit"]).sum() <= 0.004 * tot + 2
    for fld in ("nb_abs_equiv", "sum_1_per_ort_velocity", "sum_v_ort", "sum_1_per_velocity"):
        a, b = g.facets[fld], c.facets[fld]
        assert np.abs(a - b).sum() <= 5e-3 * np.abs(b).sum() + 1e-9, fld
    pairs = [(g.forces, c.forces), (g.textures["count_equiv"], c.textures["count_equiv"]),
             (g.profiles["count_equiv"], c.profiles["count_equiv"])]
    for a, b in pairs:
        if b.size and np.abs(b).sum() > 0:
            assert np.abs(a - b).sum() <= 5e-3 * np.abs(b).sum() + 2
    kg, kc_ = int(g.log_count[0]), int(c.log_count[0])
    if kc_:
        assert abs(kg - kc_) <= 2e-3 * kc_ + 2
        vg = np.sort(g.log["velocity"][:kg])
        vc = np.sort(c.log["velocity"][:kc_])
        k = min(kg, kc_)
        assert np.abs(vg[:k] - vc[:k]).mean() <= 1e-3 * vc.mean()


def test_gpu_forces_balance_in_a_closed_box(gpu):
    """The momentum balance of tests/test_oracle_features.py on the GPU's
    own tallies: the walls' summed impulse and impulse momentum equal the
    source's desorption records to rounding."""
    m = FM.force_box(torque_ref=(1.0, 2.0, 3.0))
    n = 200_000
    loaded(gpu, m, seed=8)
    gpu.run(n, max_events=MAXEV)
    t = gpu.read()
    F = t.forces[0]
    src = m.ids["source"]
    walls = [i for i in range(m.n_facets) if i != src]
    assert np.allclose(F[walls, 0:3].sum(0), F[src, 0:3], rtol=0, atol=1e-9 * np.abs(F[walls, 0:3]).sum())
    assert np.allclose(F[walls, 6:9].sum(0), F[src, 6:9], rtol=0, atol=1e-9 * np.abs(F[walls, 6:9]).sum())
    assert t.glob[0]["nb_abs_equiv"] == n


def test_gpu_moving_floor_identity(gpu):
    """|v_out - v_wall| equals the bounce speed for every logged particle."""
    m = FEATURES["translation"]()
    loaded(gpu, m, seed=4)
    gpu.run(20_000, max_events=MAXEV)
    t = gpu.read()
    k = int(t.log_count[0])
    L = t.log[:k]
    cm = CpuModel(m)
    O, U, V, Nn = cm.frame(m.ids["ceiling"])
    nU, nV = U / np.linalg.norm(U), V / np.linalg.norm(V)
    th, ph = L["hit_theta"], L["hit_phi"]
    d = (np.sin(th) * np.cos(ph))[:, None] * nU + (np.sin(th) * np.sin(ph))[:, None] * nV + np.cos(th)[:, None] * Nn
    vv = L["velocity"][:, None] * d - np.array([25.0, -10.0, 0.0])
    v1 = 145.469 * math.sqrt(293.15) / 28.0
    assert k > 19_000
    assert np.allclose(np.linalg.norm(vv, axis=1), v1, rtol=1e-9)


def test_gpu_accommodation_speed_is_exact(gpu):
    m = FEATURES["accommodation"]()
    loaded(gpu, m, seed=6)
    gpu.run(5000, max_events=MAXEV)
    t = gpu.read()
    v = t.log["velocity"][: int(t.log_count[0])]
    v0 = FM.v_nonmaxwell_desorb(300.0, 28.0)
    v1 = math.sqrt(v0 ** 2 + (29369.939 * 600.0 / 28.0 - v0 ** 2) * 0.5)
    assert len(v) > 4990 and np.allclose(v, v1, rtol=1e-13)


def test_gpu_joined_tubes_transmit_like_one_tube(gpu):
    """Config-1 sized run of the teleport-joined tubes against the single
    tube on the GPU: the same transmission within 3 sigma."""
    from molflow_sources_amd import geometry
    n = 2_000_000
    joined = FM.teleport_tubes(rotate=True)
    loaded(gpu, joined, seed=3)
    gpu.run(n, max_events=MAXEV)
    tj = gpu.read()
    loaded(gpu, geometry.tube_model(), seed=3)
    gpu.run(n, max_events=MAXEV)
    ts = gpu.read()
    pj = tj.facets[0]["nb_abs_equiv"][joined.ids["B_out"]] / n
    ps = ts.facets[0]["nb_abs_equiv"][1] / n
    assert int(tj.glob[0]["nb_leak_total"]) == 0
    assert abs(pj - ps) < 3 * math.sqrt(2 * ps * (1 - ps) / n)
