"""Small models that each exercise one branch of the tracer which the
reference's fixtures leave open (specular / cos^N / mixed reflection,
partial accommodation, moving facets, forces, hard-sphere scattering,
outgassing maps, angular and ortho/tangential velocity profiles, teleports).
Every model is built so that a closed form or an exact identity predicts
what it records; tests/test_oracle_features.py checks the oracle against
those predictions and tests/test_gpu_features.py checks the HIP tracer
against the oracle on the same seeds.
"""
from __future__ import annotations

import math

import numpy as np

from molflow_sources_amd import _abi as A
from molflow_sources_amd.model import Model, default_facets, default_settings

KB = 1.38e-23          # the reference's Boltzmann constant (Particle.cpp:427)
AMU = 1.67e-27         # CDFGeneration.cpp:60


def newell(p: np.ndarray) -> np.ndarray:
    q = np.roll(p, -1, axis=0)
    return np.array([np.sum((p[:, 1] - q[:, 1]) * (p[:, 2] + q[:, 2])),
                     np.sum((p[:, 2] - q[:, 2]) * (p[:, 0] + q[:, 0])),
                     np.sum((p[:, 0] - q[:, 0]) * (p[:, 1] + q[:, 1]))])


def facing(points, normal) -> np.ndarray:
    """The polygon's vertices ordered so that its Newell normal points along
    `normal` (the side a one-sided facet is hit from)."""
    p = np.asarray(points, dtype=np.float64)
    return p if np.dot(newell(p), normal) > 0 else p[::-1].copy()


class Builder:
    """Collects polygons (each with its own vertices) and facet settings."""

    def __init__(self):
        self.verts: list[np.ndarray] = []
        self.polys: list[list[int]] = []
        self.params: list[dict] = []
        self.names: dict[str, int] = {}

    def add(self, name: str, points, normal, **fields) -> int:
        p = facing(points, normal)
        base = sum(len(v) for v in self.verts)
        self.verts.append(p)
        self.polys.append(list(range(base, base + len(p))))
        self.params.append(fields)
        self.names[name] = len(self.polys) - 1
        return self.names[name]

    def model(self, settings=None, name="feature", n_structures=1) -> Model:
        F = default_facets(len(self.polys))
        for i, kv in enumerate(self.params):
            for k, v in kv.items():
                F[k][i] = v
        m = Model(np.concatenate(self.verts), self.polys, F, settings=settings or default_settings(),
                  n_structures=n_structures, name=name)
        m.ids = dict(self.names)
        return m


def square(cx, cy, z, half):
    return [[cx - half, cy - half, z], [cx + half, cy - half, z], [cx + half, cy + half, z], [cx - half, cy + half, z]]


def thin_box(floor: dict, source: dict | None = None, ceiling: dict | None = None, settings=None,
             width: float = 2000.0, height: float = 1.0, source_z: float = 0.5, source_half: float = 0.5,
             log_limit: int = 0) -> Model:
    """A flat box, `width` wide and `height` tall: the floor (z = 0, facing
    up) carries the reflection under test; the ceiling (z = height, facing
    down) absorbs every particle the floor sends up and logs it; the four
    side walls absorb. The source is a small square at source_z facing down
    (cosine desorption unless `source` says otherwise) with opacity 0, so a
    particle the floor reflects passes through it unchanged. Every particle
    therefore meets the floor once and the ceiling's log records the
    reflected direction and speed of one floor bounce."""
    b = Builder()
    w = width / 2.0
    src = dict(desorb_type=A.DES_COSINE, outgassing=1.0, opacity=0.0, sticking=0.0)
    src.update(source or {})
    b.add("source", square(0.0, 0.0, source_z, source_half), [0, 0, -1], **src)
    fl = dict(sticking=0.0)
    fl.update(floor)
    b.add("floor", square(0.0, 0.0, 0.0, w), [0, 0, 1], **fl)
    ce = dict(sticking=1.0)
    ce.update(ceiling or {})
    b.add("ceiling", square(0.0, 0.0, height, w), [0, 0, -1], **ce)
    for k, (nx, ny) in enumerate(((1, 0), (-1, 0), (0, 1), (0, -1))):
        # side wall at x = -w (normal +x) etc.
        if nx:
            x = -w * nx
            pts = [[x, -w, 0], [x, w, 0], [x, w, height], [x, -w, height]]
        else:
            y = -w * ny
            pts = [[-w, y, 0], [w, y, 0], [w, y, height], [-w, y, height]]
        b.add(f"side{k}", pts, [nx, ny, 0], sticking=1.0)
    s = settings or default_settings()
    if log_limit:
        s.log_enabled = 1
        s.log_facet_id = b.names["ceiling"]
        s.log_limit = log_limit
    return b.model(s, name="thin_box")


def cube(edge: float = 10.0) -> Builder:
    """Six inward-facing faces of [0, edge]^3 (names xlo, xhi, ylo, ...)."""
    b = Builder()
    e = edge
    faces = {
        "zlo": ([[0, 0, 0], [e, 0, 0], [e, e, 0], [0, e, 0]], [0, 0, 1]),
        "zhi": ([[0, 0, e], [e, 0, e], [e, e, e], [0, e, e]], [0, 0, -1]),
        "xlo": ([[0, 0, 0], [0, e, 0], [0, e, e], [0, 0, e]], [1, 0, 0]),
        "xhi": ([[e, 0, 0], [e, e, 0], [e, e, e], [e, 0, e]], [-1, 0, 0]),
        "ylo": ([[0, 0, 0], [e, 0, 0], [e, 0, e], [0, 0, e]], [0, 1, 0]),
        "yhi": ([[0, e, 0], [e, e, 0], [e, e, e], [0, e, e]], [0, -1, 0]),
    }
    for name, (pts, n) in faces.items():
        b.add(name, pts, n, sticking=0.0)
    return b


def force_box(torque_ref=(1.0, 2.0, 3.0), absorber_sticking: float = 0.3) -> Model:
    """Closed cube with diffuse walls (one of them partly absorbing) and a
    tiny transparent source at (5, 5, 5) facing -z, forces on. The source's
    impulse records are its desorptions only (later passes through a
    transparent facet carry no force terms), so the walls together must
    record the same total impulse and impulse momentum (see the tests)."""
    b = cube(10.0)
    b.params[b.names["xhi"]]["sticking"] = absorber_sticking
    b.add("source", square(5.0, 5.0, 5.0, 0.005), [0, 0, -1], desorb_type=A.DES_COSINE, outgassing=1.0,
          opacity=0.0)
    s = default_settings()
    s.enable_force = 1
    s.use_maxwell = 0
    for k in range(3):
        s.torque_ref[k] = torque_ref[k]
    return b.model(s, name="force_box")


def equilibrium_box(pump_half: float = 0.5) -> Model:
    """Closed isothermal cube with diffuse walls, a small two-sided pump
    (s = 1) floating 2 cm above the floor and a tiny transparent source in
    the middle: the walls see gas close to equilibrium, so their pressure,
    density, impingement rate and mean speed obey P = n k T and
    Z = n <v> / 4."""
    b = cube(10.0)
    b.add("pump", square(5.0, 5.0, 2.0, pump_half), [0, 0, 1], sticking=1.0, is_2sided=1)
    b.add("source", square(5.0, 5.0, 5.0, 0.005), [0, 0, -1], desorb_type=A.DES_COSINE, outgassing=1.0,
          opacity=0.0)
    return b.model(default_settings(), name="equilibrium_box")


def scatter_beam(mfp: float = 2.0, mass_ratio: float = 1.0, rho: float = 1.0, cutoff: float = 0.0) -> Model:
    """The thin box with an absorbing floor, a collimated source (cos^N,
    N = 1e5) 1 cm above it and background-gas scattering of mean free path
    `mfp` cm."""
    s = default_settings()
    s.scatter_enabled = 1
    s.scatter_mean_free_path_cm = mfp
    s.scatter_mass_ratio = mass_ratio
    s.scatter_rho = rho
    if cutoff > 0:
        s.scatter_enable_cutoff = 1
        s.scatter_cutoff_speed = cutoff
    s.use_maxwell = 0
    return thin_box(floor=dict(sticking=1.0), source=dict(desorb_type=A.DES_COSINE_N, desorb_exponent=1e5),
                    settings=s, height=3.0, source_z=1.0, width=200.0)


def og_map_plate(ratio: float = 1.3, seed: int = 4):
    """A 4 cm x 3 cm plate desorbing from an outgassing map of
    ceil(4 ratio) x ceil(3 ratio) cells (random values, one cell zero) in an
    absorbing box, with a texture of the same cells counting desorptions.
    Returns the model and the map."""
    b = Builder()
    wu, wv = 4.0, 3.0
    plate = [[0, 0, 0], [wu, 0, 0], [wu, wv, 0], [0, wv, 0]]
    mw, mh = math.ceil(wu * ratio), math.ceil(wv * ratio)
    rng = np.random.default_rng(seed)
    og = rng.uniform(0.2, 1.0, (mh, mw)) * 1e-3
    og[1, 2] = 0.0
    b.add("plate", plate, [0, 0, 1], desorb_type=A.DES_COSINE, use_outgassing_map=1, og_map_width=mw,
          og_map_height=mh, og_ratio_u=ratio, og_ratio_v=ratio, total_outgassing=float(og.sum()),
          is_textured=1, count_des=1, tex_width_precise=wu * ratio, tex_height_precise=wv * ratio, sticking=1.0)
    lid = [[-1, -1, 2], [wu + 1, -1, 2], [wu + 1, wv + 1, 2], [-1, wv + 1, 2]]
    b.add("lid", lid, [0, 0, -1], sticking=1.0)
    for k, (pts, n) in enumerate((([[-1, -1, -1], [-1, wv + 1, -1], [-1, wv + 1, 2], [-1, -1, 2]], [1, 0, 0]),
                                  ([[wu + 1, -1, -1], [wu + 1, wv + 1, -1], [wu + 1, wv + 1, 2], [wu + 1, -1, 2]], [-1, 0, 0]),
                                  ([[-1, -1, -1], [wu + 1, -1, -1], [wu + 1, -1, 2], [-1, -1, 2]], [0, 1, 0]),
                                  ([[-1, wv + 1, -1], [wu + 1, wv + 1, -1], [wu + 1, wv + 1, 2], [-1, wv + 1, 2]], [0, -1, 0]),
                                  ([[-1, -1, -1], [wu + 1, -1, -1], [wu + 1, wv + 1, -1], [-1, wv + 1, -1]], [0, 0, 1]))):
        b.add(f"wall{k}", pts, n, sticking=1.0)
    m = b.model(name="og_map_plate")
    m.og_maps[b.names["plate"]] = og
    return m, og


def _ring(n, r, z):
    a = np.arange(n) / n * 2.0 * math.pi
    return np.stack([r * np.cos(a), r * np.sin(a), np.full(n, z)], axis=1)


def _half_tube(b: Builder, tag: str, length: float, n_az: int, n_ax: int, R: np.ndarray, t: np.ndarray,
               start: tuple, end: tuple):
    """Walls of a tube along local +z (radius 1) and its two end disks, moved
    by x -> R x + t. start / end are (name, local normal z sign, fields)."""
    move = lambda p: np.asarray(p, dtype=np.float64) @ R.T + t  # noqa: E731
    rings = [_ring(n_az, 1.0, length * k / n_ax) for k in range(n_ax + 1)]
    for k in range(n_ax):
        for j in range(n_az):
            j1 = (j + 1) % n_az
            q = np.array([rings[k][j], rings[k + 1][j], rings[k + 1][j1], rings[k][j1]])
            c = q.mean(0)
            b.add(f"{tag}_w{k}_{j}", move(q), R @ np.array([-c[0], -c[1], 0.0]), sticking=0.0)
    for (name, sign, fields), ring in ((start, rings[0]), (end, rings[-1])):
        b.add(name, move(ring), R @ np.array([0.0, 0.0, float(sign)]), **fields)


def teleport_tubes(rotate: bool = True, n_az: int = 64, n_ax: int = 25):
    """Two tubes of radius 1 cm and length 5 cm joined by a teleport pair:
    tube A's far disk sends particles to tube B's near disk and back. Tube A
    starts with the cosine source disk (absorbing on return, like the
    reference's pipe), tube B ends with an absorbing disk. B sits 20 cm
    away and, with rotate, is turned 90 degrees about y, so the teleport
    must map directions between different facet frames. The pair is
    equivalent to one 10 cm tube with 2 x n_ax rings."""
    b = Builder()
    R = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]]) if rotate else np.eye(3)
    src = dict(desorb_type=A.DES_COSINE, outgassing=1.0, sticking=1.0)
    # facet ids are fixed by insertion order: A walls, A_in, A_out, B walls, B_in, B_out
    nA = n_az * n_ax
    a_out = nA + 1
    b_in = 2 * nA + 2
    _half_tube(b, "A", 5.0, n_az, n_ax, np.eye(3), np.zeros(3), ("A_in", 1, src),
               ("A_out", -1, dict(is_2sided=1, teleport_dest=b_in + 1)))
    _half_tube(b, "B", 5.0, n_az, n_ax, R, np.array([20.0, 0.0, 0.0]),
               ("B_in", -1, dict(is_2sided=1, teleport_dest=a_out + 1)), ("B_out", -1, dict(sticking=1.0)))
    m = b.model(name="teleport_tubes")
    assert m.ids["A_out"] == a_out and m.ids["B_in"] == b_in
    return m


def teleport_back_chain():
    """Two closed boxes 100 cm apart along x. Box 1 (around the source)
    has wall X (its +x wall) teleporting to wall Y of box 2 (its -x wall);
    box 2's wall Z (its +y wall) teleports "back" (destination -1). In the
    reference's live code a teleport records its destination as the place
    the particle came from (teleportedFrom = destIndex, Particle.cpp:122),
    so Z sends particles to Y, into box 2. Had it recorded the source X, a
    particle would reappear at X heading out of box 1 and hit the absorbing
    guard plate outside it. A teleport keeps the direction's components in
    the facet frames, so X and Z face out of their boxes (2-sided, hit from
    inside) and Y faces into box 2."""
    b = Builder()
    e = 4.0
    for box, ox in ((1, 0.0), (2, 100.0)):
        faces = {
            "zlo": ([[0, 0, 0], [e, 0, 0], [e, e, 0], [0, e, 0]], [0, 0, 1]),
            "zhi": ([[0, 0, e], [e, 0, e], [e, e, e], [0, e, e]], [0, 0, -1]),
            "xlo": ([[0, 0, 0], [0, e, 0], [0, e, e], [0, 0, e]], [1, 0, 0]),
            "xhi": ([[e, 0, 0], [e, e, 0], [e, e, e], [e, 0, e]], [-1, 0, 0]),
            "ylo": ([[0, 0, 0], [e, 0, 0], [e, 0, e], [0, 0, e]], [0, 1, 0]),
            "yhi": ([[0, e, 0], [e, e, 0], [e, e, e], [0, e, e]], [0, -1, 0]),
        }
        if box == 1:
            faces["xhi"] = (faces["xhi"][0], [1, 0, 0])     # X faces out of box 1
        else:
            faces["yhi"] = (faces["yhi"][0], [0, 1, 0])     # Z faces out of box 2
        for name, (pts, n) in faces.items():
            p = np.asarray(pts, dtype=np.float64) + np.array([ox, 0, 0])
            b.add(f"b{box}_{name}", p, n, sticking=0.0)
    ids = b.names
    b.params[ids["b1_xhi"]].update(teleport_dest=ids["b2_xlo"] + 1, is_2sided=1)   # X -> Y
    b.params[ids["b1_zlo"]].update(sticking=0.2)
    b.params[ids["b2_yhi"]].update(teleport_dest=-1, is_2sided=1)                   # Z -> back
    b.params[ids["b2_zhi"]].update(sticking=0.2)
    b.add("b1_guard", [[e + 1, -1, -1], [e + 1, e + 1, -1], [e + 1, e + 1, e + 1], [e + 1, -1, e + 1]],
          [-1, 0, 0], sticking=1.0)
    b.add("source", square(2.0, 2.0, 2.0, 0.5), [0, 0, -1], desorb_type=A.DES_COSINE, outgassing=1.0, opacity=0.0)
    return b.model(name="teleport_back_chain")


def with_settings(**kw):
    s = default_settings()
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return s


def v_nonmaxwell_desorb(T: float, m: float) -> float:
    """Desorption speed without Maxwell sampling (Particle.cpp:483)."""
    return 145.469 * math.sqrt(T / m)


def mean_v2_wall_flux(T: float, m: float) -> float:
    """E[v^2] of the wall-collision Maxwell law that CDFGeneration.cpp:55-83
    tabulates, P(v < x) = 1 - exp(-q)(1 + q), q = x^2/(2a^2), a^2 = kT/m:
    v^2/(2a^2) is Gamma(2, 1), so E[v^2] = 4 a^2."""
    return 4.0 * KB * T / (m * AMU)
