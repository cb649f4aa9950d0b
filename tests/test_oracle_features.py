"""Pins of the oracle's branches that the reference's fixtures leave open.

Each test runs the oracle (oracle/mfcpu.cpp) on a small model from
tests/feature_models.py and checks what it records against a closed form
or an exact identity that follows from the reference's physics, not
against a retyped copy of the formula:

  reflection (Particle.cpp:779-801)  diffuse P(cos<=x)=x^2, cos^N x^(N+1),
                                     specular keeps the angle and mirrors the
                                     path, mixtures add with their weights
  UpdateVelocity (:1138-1163)        exact speeds without Maxwell sampling,
                                     E[v^2] relaxation with it
  TreatMovingFacet (Physics.cpp:22-39) |v_out - v_wall| = v_bounce per particle
  forces (:1208-1212, :659-675)      momentum and angular-momentum balance
  PerformScatter (:840-886)          exponential free path, deflection law
  outgassing map (:425-447, :530-560) desorption positions follow the map
  ProfileFacet (:1025-1085)          angular sin^2 law, ortho / tangential
                                     speed laws
  PerformTeleport (:76-183)          joined tubes transmit like one tube;
                                     "back" returns to the last destination
"""
import math

import numpy as np
import pytest
from scipy import stats

import feature_models as FM
from molflow_sources_amd import _abi as A
from molflow_sources_amd import geometry
from molflow_sources_amd.model import default_settings
from oracle.mfcpu import CpuModel

N = 60_000


def logged(model, n=N, seed=3):
    """Run the oracle; return the ceiling log as (cos of the angle to +z,
    speed, world direction, world position) and the tallies."""
    cm = CpuModel(model)
    t = cm.run(n, seed=seed)
    k = int(t.log_count[0])
    L = t.log[:k]
    O, U, V, Nn = cm.frame(model.ids["ceiling"])
    nU, nV = U / np.linalg.norm(U), V / np.linalg.norm(V)
    th, ph = L["hit_theta"], L["hit_phi"]
    d = (np.sin(th) * np.cos(ph))[:, None] * nU + (np.sin(th) * np.sin(ph))[:, None] * nV + np.cos(th)[:, None] * Nn
    pos = O + L["u"][:, None] * U + L["v"][:, None] * V
    return d[:, 2], L["velocity"].copy(), d, pos, t


def ks_p(sample, cdf):
    return stats.kstest(sample, cdf).pvalue


# ------------------------------------------------------------- reflection laws
def test_diffuse_reflection_follows_the_cosine_law(oracle_lib):
    m = FM.thin_box(floor=dict(), log_limit=N)
    c, v, d, pos, t = logged(m)
    assert len(c) > 0.999 * N and int(t.glob[0]["nb_leak_total"]) == 0
    assert ks_p(c, lambda x: np.clip(x, 0, 1) ** 2) > 1e-3
    assert abs(c.mean() - 2 / 3) < 4 * math.sqrt(1 / 18 / len(c))


@pytest.mark.parametrize("n_exp", [3.0, 10.0])
def test_cos_n_reflection_law(oracle_lib, n_exp):
    m = FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=0.0, cosine_exponent=n_exp), log_limit=N)
    c, *_ = logged(m)
    assert ks_p(c, lambda x: np.clip(x, 0, 1) ** (n_exp + 1)) > 1e-3
    # E[cos] = (N+1)/(N+2)
    var = (n_exp + 1) / (n_exp + 3) - ((n_exp + 1) / (n_exp + 2)) ** 2
    assert abs(c.mean() - (n_exp + 1) / (n_exp + 2)) < 4 * math.sqrt(var / len(c))


def test_specular_reflection_keeps_the_angle_and_mirrors_the_path(oracle_lib):
    """A uniform source (P(cos<=x) = x) seen after a specular floor keeps
    its law (a diffuse floor would give x^2); every particle lands where the
    mirrored straight path from the source lands: its horizontal distance
    from the source centre is tan(theta) times the path height (source to
    floor, floor to ceiling), within the source's half-diagonal."""
    m = FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=1.0), source=dict(desorb_type=A.DES_UNIFORM),
                    log_limit=N)
    c, v, d, pos, t = logged(m)
    assert int(t.glob[0]["nb_leak_total"]) == 0
    assert ks_p(c, lambda x: np.clip(x, 0, 1)) > 1e-3
    keep = c > 0.05                         # skip grazing paths that leave the box sideways
    r = np.hypot(pos[keep, 0], pos[keep, 1])
    tan = np.sqrt(1 - c[keep] ** 2) / c[keep]
    assert np.all(np.abs(r - tan * (0.5 + 1.0)) <= 0.5 * math.sqrt(2) + 1e-9)
    # and the direction keeps its azimuth: the landing point lies along it
    az_pos = np.arctan2(pos[keep, 1], pos[keep, 0])
    az_dir = np.arctan2(d[keep, 1], d[keep, 0])
    far = r > 5
    assert np.all(np.abs(np.angle(np.exp(1j * (az_pos - az_dir)))[far[: len(az_pos)]]) < 0.15)


def test_mixed_reflection_adds_the_three_laws(oracle_lib):
    m = FM.thin_box(floor=dict(diffuse_part=0.3, specular_part=0.3, cosine_exponent=3.0),
                    source=dict(desorb_type=A.DES_UNIFORM), log_limit=N)
    c, *_ = logged(m)
    assert ks_p(c, lambda x: 0.3 * np.clip(x, 0, 1) ** 2 + 0.3 * np.clip(x, 0, 1) + 0.4 * np.clip(x, 0, 1) ** 4) > 1e-3
    # a mixture that forgot the specular part fails the same test
    assert ks_p(c, lambda x: 0.5 * np.clip(x, 0, 1) ** 2 + 0.5 * np.clip(x, 0, 1) ** 4) < 1e-6


# ------------------------------------------------------------- accommodation
def test_partial_accommodation_without_maxwell_is_exact(oracle_lib):
    """v1^2 = v0^2 + (29369.939 T/m - v0^2) acc for every particle
    (Particle.cpp:1153-1161), v0 = 145.469 sqrt(T_source/m)."""
    s = FM.with_settings(use_maxwell=0, gas_mass=28.0)
    m = FM.thin_box(floor=dict(accomodation=0.5, temperature=600.0), source=dict(temperature=300.0), settings=s,
                    log_limit=2000)
    _, v, *_ = logged(m, n=2000)
    v0 = FM.v_nonmaxwell_desorb(300.0, 28.0)
    v1 = math.sqrt(v0 ** 2 + (29369.939 * 600.0 / 28.0 - v0 ** 2) * 0.5)
    assert np.allclose(v, v1, rtol=1e-13)


def test_full_accommodation_without_maxwell_uses_the_reference_speed(oracle_lib):
    """Full accommodation without Maxwell sampling sets 145.469 sqrt(T)/m
    (Particle.cpp:1152, sqrt(T)/m rather than sqrt(T/m))."""
    s = FM.with_settings(use_maxwell=0, gas_mass=4.0)
    m = FM.thin_box(floor=dict(temperature=500.0), settings=s, log_limit=500)
    _, v, *_ = logged(m, n=500)
    assert np.allclose(v, 145.469 * math.sqrt(500.0) / 4.0, rtol=1e-13)


def test_partial_accommodation_with_maxwell_relaxes_the_mean_square(oracle_lib):
    """E[v1^2] = (1-acc) E[v0^2] + acc E[v_wall^2], both wall-flux Maxwell
    laws: E[v^2] = 4kT/m."""
    s = FM.with_settings(use_maxwell=1, gas_mass=28.0)
    acc = 0.3
    m = FM.thin_box(floor=dict(accomodation=acc, temperature=900.0), source=dict(temperature=300.0), settings=s,
                    log_limit=N)
    _, v, *_ = logged(m)
    want = (1 - acc) * FM.mean_v2_wall_flux(300.0, 28.0) + acc * FM.mean_v2_wall_flux(900.0, 28.0)
    v2 = v ** 2
    assert abs(v2.mean() - want) < 4 * v2.std() / math.sqrt(len(v2)) + 2e-3 * want


# ------------------------------------------------------------- moving facets
def test_translating_floor_adds_its_velocity(oracle_lib):
    """v_out = v1 d + a with |d| = 1 and v1 the bounce speed: every logged
    velocity vector minus the wall velocity has length v1."""
    s = FM.with_settings(use_maxwell=0, gas_mass=28.0, motion_type=1, motion_vector2=(25.0, -10.0, 0.0))
    m = FM.thin_box(floor=dict(is_moving=1), settings=s, log_limit=20_000)
    _, v, d, pos, t = logged(m, n=20_000)
    v1 = 145.469 * math.sqrt(293.15) / 28.0
    vv = v[:, None] * d - np.array([25.0, -10.0, 0.0])
    assert np.allclose(np.linalg.norm(vv, axis=1), v1, rtol=1e-9)
    assert np.all(vv[:, 2] > 0)


def test_rotating_floor_adds_omega_cross_r(oracle_lib):
    """Rotation about base point b with vector w adds w x 0.01 (r - b) at
    the bounce point r (found by running the logged path back to z = 0)."""
    base, w = np.array([30.0, -20.0, 0.0]), np.array([0.0, 0.0, 400.0])
    s = FM.with_settings(use_maxwell=0, gas_mass=28.0, motion_type=2, motion_vector1=tuple(base),
                         motion_vector2=tuple(w))
    m = FM.thin_box(floor=dict(is_moving=1), settings=s, log_limit=20_000)
    _, v, d, pos, t = logged(m, n=20_000)
    r = pos - (pos[:, 2] / d[:, 2])[:, None] * d
    a = np.cross(w, 0.01 * (r - base))
    v1 = 145.469 * math.sqrt(293.15) / 28.0
    vv = v[:, None] * d - a
    assert np.allclose(np.linalg.norm(vv, axis=1), v1, rtol=1e-7)


# ------------------------------------------------------------- forces
def test_forces_balance_in_a_closed_box(oracle_lib):
    """A particle desorbs with momentum p0 at r0 (+p0 and r0 x p0 on the
    source), each bounce records p_in - p_out, the absorption p_last. Along
    a straight flight r_(k+1) - r_k is parallel to p_k, so the walls'
    records telescope to p0 and r0 x p0 exactly: the walls together equal
    the source, per component, to rounding. The torque reference point
    shifts both sides alike; the source being tiny, its impulse momentum is
    (centre - ref) x its impulse."""
    ref = np.array([1.0, 2.0, 3.0])
    m = FM.force_box(torque_ref=tuple(ref))
    n = 20_000
    t = CpuModel(m).run(n, seed=2)
    F = t.forces[0]                      # [facet][9]
    src = m.ids["source"]
    walls = [i for i in range(m.n_facets) if i != src]
    imp_src, imp_w = F[src, 0:3], F[walls, 0:3].sum(0)
    mom_src, mom_w = F[src, 6:9], F[walls, 6:9].sum(0)
    scale = np.abs(F[walls, 0:3]).sum()
    assert np.allclose(imp_w, imp_src, rtol=0, atol=1e-10 * scale)
    assert np.allclose(mom_w, mom_src, rtol=0, atol=1e-10 * np.abs(F[walls, 6:9]).sum())
    assert np.allclose(mom_src, np.cross(np.array([5.0, 5.0, 5.0]) - ref, imp_src), rtol=2e-3,
                       atol=2e-3 * np.abs(mom_src).max())
    # impulse squares of a cosine source of speed v0, normal -z:
    # E[vz^2] = v0^2/2, E[vx^2] = E[vy^2] = v0^2/4
    v0 = FM.v_nonmaxwell_desorb(293.15, 28.0)
    sq = F[src, 3:6] / (n * v0 ** 2)
    assert abs(sq[2] - 0.5) < 0.01 and abs(sq[0] - 0.25) < 0.01 and abs(sq[1] - 0.25) < 0.01
    # and every desorbed particle was absorbed (closed box)
    assert t.glob[0]["nb_abs_equiv"] == n


# ------------------------------------------------------------- scattering
def test_scatter_free_path_is_exponential_with_the_mean_free_path(oracle_lib):
    """Collimated beam 1 cm above an absorbing floor: the first event after
    desorption is the floor with probability exp(-1 / mfp), else a scatter."""
    mfp = 2.0
    m = FM.scatter_beam(mfp=mfp)
    n = 40_000
    f, k = CpuModel(m).trace(n, 3, seed=7)
    first = k[:, 1]
    assert np.all((first == A.EV_ABSORB) | (first == A.EV_SCATTER))
    p = math.exp(-1.0 / mfp)
    got = (first == A.EV_ABSORB).mean()
    assert abs(got - p) < 4 * math.sqrt(p * (1 - p) / n)


def test_scatter_deflection_and_speed_laws(oracle_lib):
    """Physics.cpp:41-63 with b = rho sqrt(u): for equal masses the
    deflection has cos(theta) = sqrt(u), so P(cos <= x) = x^2, and the new
    speed is v cos(theta) exactly (elastic equal-mass collision); for an
    infinitely heavy background (mass ratio 0) the speed is kept and, with
    atan's range, cos(theta) = |1 - 2u| is uniform on [0, 1]."""
    from oracle.mfcpu import scatter_sample
    c, v = scatter_sample(1.0, 1.0, 50_000, seed=3)
    assert ks_p(c, lambda x: np.clip(x, 0, 1) ** 2) > 1e-3
    assert np.allclose(v, c, rtol=1e-9, atol=1e-12)
    c0, v0 = scatter_sample(0.0, 2.5, 50_000, seed=4)
    assert np.allclose(v0, 1.0, rtol=1e-15)
    assert ks_p(c0, lambda x: np.clip(x, 0, 1)) > 1e-3
    # a heavier test particle is deflected less
    c4, _ = scatter_sample(4.0, 1.0, 20_000, seed=5)
    assert c4.mean() > c.mean() + 0.05


# ------------------------------------------------------------- outgassing map
def test_outgassing_map_places_desorptions_by_the_map(oracle_lib):
    """Desorption cell counts (a texture of the map's cells counting
    desorptions) follow the map values, including the narrower last column
    and row of a map whose size is not a whole number of cells; the empty
    cell gets none."""
    m, og = FM.og_map_plate(ratio=1.3)
    n = 60_000
    cm = CpuModel(m)
    t = cm.run(n, seed=5)
    h, w = og.shape
    tex = t.texture_of(m.ids["plate"], w, h)["count_equiv"]
    assert tex.sum() == n
    assert tex[1, 2] == 0
    p = (og / og.sum()).ravel()
    ok = p > 0
    chi2 = stats.chisquare(tex.ravel()[ok], n * p[ok])
    assert chi2.pvalue > 1e-3
    # a desorption map that ignored the cell widths would put the narrow
    # last column's share at a full cell's
    assert t.facets[0]["nb_desorbed"][m.ids["plate"]] == n


# ------------------------------------------------------------- profiles
def test_angular_profile_of_a_cosine_flux(oracle_lib):
    """The ceiling sees the diffuse floor's cosine law: P(theta < t) =
    sin^2 t, so slice i of the angular profile (theta in 100 slices of
    [0, pi/2]) gets sin^2 t_(i+1) - sin^2 t_i."""
    m = FM.thin_box(floor=dict(), ceiling=dict(profile_type=A.PROFILE_ANGULAR))
    n = 80_000
    t = CpuModel(m).run(n, seed=9)
    prof = t.profile_of(m.ids["ceiling"])["count_equiv"]
    edges = np.linspace(0, math.pi / 2, 101)
    p = np.diff(np.sin(edges) ** 2)
    assert abs(prof.sum() - n) < 1e-3 * n
    exp_ = prof.sum() * p
    res = stats.chisquare(prof, exp_ * prof.sum() / exp_.sum())
    assert res.pvalue > 1e-3


@pytest.mark.parametrize("ptype", [A.PROFILE_ORT_VELOCITY, A.PROFILE_TAN_VELOCITY, A.PROFILE_VELOCITY])
def test_ortho_and_tangential_velocity_profiles(oracle_lib, ptype):
    """Fixed bounce speed v1 (no Maxwell sampling, partial accommodation
    from a hot floor) and a cos^3 floor: the ceiling's orthogonal speed
    v1 cos(theta) has P(cos <= x) = x^4, the tangential v1 sin(theta) has
    P(sin <= y) = 1 - (1 - y^2)^2, the speed v1 all in one slice. Slices
    are 1/100 of 4 sqrt(2 8.31 T / 0.001 / m), T the ceiling's."""
    s = FM.with_settings(use_maxwell=0, gas_mass=28.0)
    m = FM.thin_box(floor=dict(diffuse_part=0.0, specular_part=0.0, cosine_exponent=3.0, accomodation=0.99,
                               temperature=10_000.0),
                    ceiling=dict(profile_type=ptype, temperature=1324.0), settings=s)
    n = 60_000
    t = CpuModel(m).run(n, seed=12)
    prof = t.profile_of(m.ids["ceiling"])["count_equiv"]
    v0 = FM.v_nonmaxwell_desorb(293.15, 28.0)
    v1 = math.sqrt(v0 ** 2 + (29369.939 * 10_000.0 / 28.0 - v0 ** 2) * 0.99)
    vmax = 4.0 * math.sqrt(2.0 * 8.31 * 1324.0 / 0.001 / 28.0)
    x = np.minimum(np.arange(101) * vmax / 100.0 / v1, 1.0)
    if ptype == A.PROFILE_VELOCITY:
        k = int(v1 / vmax * 100)
        assert prof[k] == pytest.approx(prof.sum()) and prof.sum() > 0.999 * n
        return
    cdf = x ** 4 if ptype == A.PROFILE_ORT_VELOCITY else 1 - (1 - x ** 2) ** 2
    p = np.diff(cdf)
    ok = p * prof.sum() > 5
    res = stats.chisquare(prof[ok], prof[ok].sum() * p[ok] / p[ok].sum())
    assert res.pvalue > 1e-3


# ------------------------------------------------------------- teleports
def test_joined_tubes_transmit_like_one_tube(oracle_lib):
    """Two 5 cm tubes joined by a teleport pair (the second one turned by
    90 degrees) are one 10 cm tube for every particle: the teleport maps the
    hit point and direction into the other facet's frame. Transmission
    equals the single tube's (same seeds, so also particle by particle for
    nearly all particles) and the reference's stored L/R = 10 value."""
    joined = FM.teleport_tubes(rotate=True)
    single = geometry.tube_model()
    n = 100_000
    tj = CpuModel(joined).run(n, seed=31)
    ts = CpuModel(single).run(n, seed=31)
    pj = tj.facets[0]["nb_abs_equiv"][joined.ids["B_out"]] / n
    ps = ts.facets[0]["nb_abs_equiv"][1] / n
    assert int(tj.glob[0]["nb_leak_total"]) == 0
    assert abs(pj - ps) < 3 * math.sqrt(2 * ps * (1 - ps) / n)
    # B02 (L/R = 10, 1000-gon) stored 0.190947; the 64-gon tube is within 1 %
    assert abs(pj - 0.190947) < 0.01 * 0.190947 + 4 * math.sqrt(ps * (1 - ps) / n)
    # the same particles end in the same places: inlet-return counts agree
    assert abs(tj.facets[0]["nb_abs_equiv"][joined.ids["A_in"]] - ts.facets[0]["nb_abs_equiv"][0]) < 0.01 * n
    # without the 90-degree turn, too
    tu = CpuModel(FM.teleport_tubes(rotate=False)).run(n, seed=31)
    pu = tu.facets[0]["nb_abs_equiv"][joined.ids["B_out"]] / n
    assert abs(pu - ps) < 3 * math.sqrt(2 * ps * (1 - ps) / n)


def test_teleport_back_returns_to_the_last_destination(oracle_lib):
    """Box 1's wall X sends particles to box 2's wall Y; box 2's wall Z
    teleports "back" (-1). The event after a teleport from Z is on a wall
    of box 2 other than Y, never in box 1 or on the guard plate outside X:
    the return target is the last destination Y (teleportedFrom =
    destIndex, Particle.cpp:122)."""
    m = FM.teleport_back_chain()
    ids = m.ids
    box1 = {ids[k] for k in ids if k.startswith("b1_")} | {ids["b2_xlo"]}
    n, k = 4000, 60
    f, kd = CpuModel(m).trace(n, k, seed=2)
    z = ids["b2_yhi"]
    seen = 0
    for row in range(n):
        for j in range(k - 1):
            if kd[row, j] == A.EV_TELEPORT and f[row, j] == z and kd[row, j + 1] >= 0:
                seen += 1
                assert f[row, j + 1] not in box1, (row, j, f[row, : j + 2], kd[row, : j + 2])
    assert seen > 100


def test_adjoint_cutoff_term_matches_paired_differences(oracle_lib):
    """The low-flux cutoff's term of the sticking gradient. A closed cube
    with a transparent source and two sticky walls (groups s = 0.2 and
    0.05), cutoff 1e-3, two targets: the gradient of up . J0 against
    Richardson-extrapolated central differences (h and 2h, h = 1e-2 s_g) of
    the same particles. Without the term the path derivatives miss about
    0.7 % of the gradient here (measured with it switched off); with it the
    difference is below 0.25 % (the renewal approximation behind the term
    leaves a few 1e-4 for a group with only a few distinct weight steps)."""
    from oracle.mfcpu import CpuModel
    b = FM.cube(10.0)
    b.add("source", FM.square(5.0, 5.0, 5.0, 0.5), [0, 0, -1], desorb_type=A.DES_COSINE, outgassing=1.0,
          opacity=0.0)
    s = default_settings()
    s.low_flux_mode = 1
    s.low_flux_cutoff = 1e-3
    m = b.model(s)
    grp = np.full(m.n_facets, -1, dtype=np.int32)
    grp[m.ids["xlo"]] = 0
    grp[m.ids["ylo"]] = 1
    base = np.array([0.2, 0.05])
    for k, v in enumerate(base):
        m.facets["sticking"][grp == k] = v
    tg = np.array([m.ids["zhi"], m.ids["xhi"]], dtype=np.int32)
    cm = CpuModel(m)
    up = np.array([1.0, 0.7])
    per, nb = 20_000, 8

    def objective(sv, blk):
        j0, jac = cm.adjoint(per, grp, sv, tg, seed=9, first=blk * per)
        return float(up @ j0) / per, (up @ jac) / per

    adj = np.array([objective(base, blk)[1] for blk in range(nb)])
    fd = np.zeros_like(adj)
    for g in range(2):
        h = 1e-2 * base[g]
        for blk in range(nb):
            f = {q: objective(base + q * h * np.eye(2)[g], blk)[0] for q in (-2, -1, 1, 2)}
            fd[blk, g] = (4 * (f[1] - f[-1]) / (2 * h) - (f[2] - f[-2]) / (4 * h)) / 3
    rel = (fd - adj).mean(0) / np.abs(adj.mean(0))
    assert np.all(np.abs(rel) < 2.5e-3), rel
