"""Synthetic models of the benchmark workloads, built in code (seeded).

  tube_model        plain steady-state tube (R 1 cm, L 10 cm, 64 x 50 wall
                    quads, cosine inlet, absorbing outlet); the pipe layout
                    follows MolflowGeometry::BuildPipe
                    (reference src/MolflowGeometry.cpp:106-178)
  beam_pipe_model   R 4 cm x 2 m triangulated pipe (~100k triangles) with
                    8 side ports (R 2 cm, 20 cm) ending in s=0.05 pump
                    disks, 6 NEG bands (s=0.01), outgassing per wall area,
                    16 textured facets (256x256) and 4 angle-map facets
  time_dependent_tube  the tube with Maxwell speeds, a 1 ms pulse and 100
                    moments of 0.5 ms over 0-50 ms, plus a transparent axial
                    recording facet with a 100-slice U profile
  variant_models    uneven beam-pipe variants for the batched launch
  adjoint_setup     32 NEG sticking groups and 64 target wall facets
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi as A
from .model import Model, Parameter, default_facets, default_settings


def _ring(n: int, r: float, z: float, phase: float = 0.0) -> np.ndarray:
    a = np.arange(n) / n * 2.0 * math.pi + phase
    return np.stack([r * np.cos(a), r * np.sin(a), np.full(n, z)], axis=1)


def tube_model(radius: float = 1.0, length: float = 10.0, n_az: int = 64, n_ax: int = 50,
               wall_sticking: float = 0.0, tex_ratio: float = 0.0) -> Model:
    """Cylinder along +z: facet 0 inlet (z=0, cosine desorption, s=1),
    facet 1 outlet (z=L, s=1), then n_ax*n_az planar wall quads with
    inward normals, ring by ring."""
    rings = [_ring(n_az, radius, length * k / n_ax) for k in range(n_ax + 1)]
    V = np.concatenate(rings)
    polys = [list(range(n_az)), [n_ax * n_az + (n_az - 1 - j) for j in range(n_az)]]
    for k in range(n_ax):
        for j in range(n_az):
            a, b = k * n_az + j, (k + 1) * n_az + j
            j1 = (j + 1) % n_az
            polys.append([a, b, (k + 1) * n_az + j1, k * n_az + j1])
    F = default_facets(len(polys))
    F["sticking"][0] = 1.0
    F["desorb_type"][0] = A.DES_COSINE
    F["outgassing"][0] = 1.0
    F["sticking"][1] = 1.0
    F["sticking"][2:] = wall_sticking
    if tex_ratio > 0:
        F["is_textured"][2:] = 1
        F["count_refl"][2:] = 1
        F["tex_width_precise"][2:] = tex_ratio * length / n_ax
        F["tex_height_precise"][2:] = tex_ratio * 2 * radius * math.sin(math.pi / n_az)
    return Model(V, polys, F, name=f"tube_R{radius}_L{length}_{n_az}x{n_ax}")


def time_dependent_tube(n_moments_end: str = "0.25e-3,0.5e-3,49.8e-3", window: float = 0.5e-3,
                        pulse: float = 1e-3, q_mbarls: float = 1.0) -> Model:
    """Config-3 model. The moment string "0.25e-3,0.5e-3,49.75e-3" would give
    99 moments with the reference's accumulating parse loop, so the end is
    set to 49.8e-3 to get 100 windows (centres 0.25 ms .. 49.75 ms)."""
    m = tube_model()
    # transparent axial recording facet (2-sided, opacity 0) through the axis
    nV = len(m.vertices)
    r = 0.999
    extra = np.array([[0.0, -r, 0.0], [0.0, -r, 10.0], [0.0, r, 10.0], [0.0, r, 0.0]])
    m.vertices = np.concatenate([m.vertices, extra])
    m.polygons.append([nV, nV + 1, nV + 2, nV + 3])
    F = default_facets(1)
    F["opacity"] = 0.0
    F["is_2sided"] = 1
    F["profile_type"] = A.PROFILE_U
    m.facets = np.concatenate([m.facets, F])
    s = default_settings()
    s.gas_mass = 28.0
    s.use_maxwell = 1
    s.calc_constant_flow = 0
    m.settings = s
    m.params.append(Parameter("pulse", [0.0, pulse, pulse * 1.000001], [q_mbarls, q_mbarls, 0.0]))
    m.facets["outgassing_param"][0] = 0
    m.facets["temperature"][:] = 293.0
    m.add_user_moments(n_moments_end, window)
    m.name = "tube_time_dependent"
    return m


def beam_pipe_model(n_az: int = 128, n_ax: int = 340, radius: float = 4.0, length: float = 200.0,
                    n_ports: int = 8, port_radius: float = 2.0, port_length: float = 20.0,
                    port_rings: int = 16, pump_sticking: float = 0.05, neg_sticking: float = 0.01,
                    tex_size: int = 256, n_textures: int = 16, n_anglemaps: int = 4,
                    outgassing_per_cm2: float = 1e-9, seed: int = 42) -> Model:
    """Config-2 beam pipe. Main-pipe quads are split into two triangles; each
    port removes a patch of the wall grid and lofts its boundary loop out to
    a circle, so the mesh is watertight."""
    dz = length / n_ax
    dphi = 2 * math.pi / n_az
    rings = [_ring(n_az, radius, dz * k) for k in range(n_ax + 1)]
    V = [np.concatenate(rings)]
    nV = (n_ax + 1) * n_az
    removed = np.zeros((n_ax, n_az), dtype=bool)
    ports = []
    half_w = int(math.ceil(math.asin(min(1.0, port_radius / radius)) / dphi))
    half_h = int(math.ceil(port_radius / dz))
    for p in range(n_ports):
        zc = length * (p + 0.5) / n_ports
        kc = int(round(zc / dz))
        jc = int(round((p % 4) * n_az / 4))
        k0, k1 = kc - half_h, kc + half_h          # quad rows k0..k1-1
        j0s = [(jc + d) % n_az for d in range(-half_w, half_w)]
        for k in range(k0, k1):
            for j in j0s:
                removed[k, j] = True
        # boundary loop of grid vertices (k,j) around the patch, consistent order
        loop = []
        js = [(jc + d) % n_az for d in range(-half_w, half_w + 1)]
        for j in js:
            loop.append((k0, j))
        for k in range(k0 + 1, k1 + 1):
            loop.append((k, js[-1]))
        for j in reversed(js[:-1]):
            loop.append((k1, j))
        for k in range(k1 - 1, k0, -1):
            loop.append((k, js[0]))
        ports.append((zc, jc * dphi, loop))
    polys, kinds = [], []   # kind: 0 wall, 1 cap, 2 port wall, 3 pump disk
    k_idx, j_idx = np.meshgrid(np.arange(n_ax), np.arange(n_az), indexing="ij")
    keep = ~removed
    a = (k_idx * n_az + j_idx)[keep]
    b = ((k_idx + 1) * n_az + j_idx)[keep]
    c = ((k_idx + 1) * n_az + (j_idx + 1) % n_az)[keep]
    d = (k_idx * n_az + (j_idx + 1) % n_az)[keep]
    wall_k = k_idx[keep]
    tri = np.empty((2 * len(a), 3), dtype=np.int64)
    tri[0::2] = np.stack([a, b, c], 1)
    tri[1::2] = np.stack([a, c, d], 1)
    polys.extend(tri.tolist())
    kinds.extend([0] * len(tri))
    tri_k = np.repeat(wall_k, 2)
    # end caps (normals into the pipe)
    polys.append(list(range(n_az)))
    kinds.append(1)
    polys.append([n_ax * n_az + (n_az - 1 - j) for j in range(n_az)])
    kinds.append(1)
    pump_ids = []
    for (zc, phic, loop) in ports:
        ed = np.array([math.cos(phic), math.sin(phic), 0.0])
        et = np.array([-math.sin(phic), math.cos(phic), 0.0])
        ez = np.array([0.0, 0.0, 1.0])
        centre = np.array([0.0, 0.0, zc]) + (radius + port_length) * ed
        base = np.array([V[0][k * n_az + j] for (k, j) in loop])
        ang = np.arctan2((base[:, 2] - zc), base @ et)
        circle = centre[None, :] + port_radius * (np.cos(ang)[:, None] * et[None, :] + np.sin(ang)[:, None] * ez[None, :])
        nl = len(loop)
        ring_ids = [np.array([k * n_az + j for (k, j) in loop])]
        for r in range(1, port_rings + 1):
            s = r / port_rings
            pts = (1 - s) * base + s * circle
            ring_ids.append(np.arange(nV, nV + nl))
            V.append(pts)
            nV += nl
        for r in range(port_rings):
            lo, hi = ring_ids[r], ring_ids[r + 1]
            for i in range(nl):
                i1 = (i + 1) % nl
                # orientation: normals point into the port (towards its axis)
                polys.append([int(lo[i]), int(lo[i1]), int(hi[i1])])
                polys.append([int(lo[i]), int(hi[i1]), int(hi[i])])
                kinds.extend([2, 2])
        pump_ids.append(len(polys))
        polys.append([int(x) for x in ring_ids[-1][::-1]])
        kinds.append(3)
    Vall = np.concatenate(V)
    m = Model(Vall, polys, default_facets(len(polys)), name=f"beam_pipe_{len(polys)}f")
    kinds = np.array(kinds)
    F = m.facets
    # fix port-wall and pump orientation by checking normals against the port axis
    _orient_ports(m, kinds, ports, radius)
    # NEG bands: 6 x 10 cm, away from the ports
    band_starts = [20.0, 45.0, 70.0, 120.0, 145.0, 170.0]
    tri_z = (tri_k + 0.5) * dz
    neg = np.zeros(len(polys), dtype=bool)
    for z0 in band_starts:
        neg[:len(tri)] |= (tri_z >= z0) & (tri_z < z0 + 10.0)
    F["sticking"][neg] = neg_sticking
    pumps = np.array(pump_ids)
    F["sticking"][pumps] = pump_sticking
    areas = m.facet_areas()
    walls = kinds != 3
    F["desorb_type"][walls] = A.DES_COSINE
    F["outgassing"][walls] = outgassing_per_cm2 * areas[walls]
    rng = np.random.default_rng(seed)
    tex_ids = list(pumps[:min(len(pumps), n_textures)])
    caps = np.nonzero(kinds == 1)[0]
    tex_ids += list(caps[:max(0, n_textures - len(tex_ids))])
    if len(tex_ids) < n_textures:
        cand = np.nonzero(neg)[0]
        tex_ids += list(rng.choice(cand, n_textures - len(tex_ids), replace=False))
    for f in tex_ids:
        F["is_textured"][f] = 1
        F["count_refl"][f] = F["count_abs"][f] = F["count_des"][f] = 1
        F["tex_width_precise"][f] = tex_size
        F["tex_height_precise"][f] = tex_size
    for f in pumps[:n_anglemaps]:
        F["am_record"][f] = 1
        F["am_phi_width"][f] = 180
        F["am_theta_lower_res"][f] = 90
        F["am_theta_higher_res"][f] = 0
        F["am_theta_limit"][f] = math.pi / 2
    m.kinds = kinds
    m.neg_mask = neg
    m.pump_ids = pumps
    m.tex_ids = np.array(tex_ids)
    m.wall_z = np.concatenate([tri_z, np.full(len(polys) - len(tri), -1.0)])
    return m


def _newell(P: np.ndarray) -> np.ndarray:
    Q = np.roll(P, -1, axis=0)
    return np.array([np.sum((P[:, 1] - Q[:, 1]) * (P[:, 2] + Q[:, 2])),
                     np.sum((P[:, 2] - Q[:, 2]) * (P[:, 0] + Q[:, 0])),
                     np.sum((P[:, 0] - Q[:, 0]) * (P[:, 1] + Q[:, 1]))])


def _orient_ports(m: Model, kinds: np.ndarray, ports, radius: float) -> None:
    """Flip port-wall triangles and pump disks whose normals point away from
    the inside of the port."""
    n_ports = len(ports)
    port_tris = np.nonzero(kinds == 2)[0]
    per_port = len(port_tris) // max(1, n_ports)
    pumps = np.nonzero(kinds == 3)[0]
    for p, (zc, phic, _) in enumerate(ports):
        ed = np.array([math.cos(phic), math.sin(phic), 0.0])
        ids = port_tris[p * per_port:(p + 1) * per_port]
        if len(ids):
            P = m.vertices[np.asarray([m.polygons[f] for f in ids], dtype=np.int64)]   # (n, 3, 3)
            c = P.mean(1)
            axis_pt = np.array([0.0, 0.0, zc]) + (c @ ed)[:, None] * ed
            Q = np.roll(P, -1, axis=1)
            nw = np.stack([np.sum((P[..., 1] - Q[..., 1]) * (P[..., 2] + Q[..., 2]), axis=1),
                           np.sum((P[..., 2] - Q[..., 2]) * (P[..., 0] + Q[..., 0]), axis=1),
                           np.sum((P[..., 0] - Q[..., 0]) * (P[..., 1] + Q[..., 1]), axis=1)], axis=1)
            for f in ids[np.einsum("ij,ij->i", nw, axis_pt - c) < 0]:
                m.polygons[f] = m.polygons[f][::-1]
        P = m.vertices[m.polygons[pumps[p]]]
        if np.dot(_newell(P), -ed) < 0:
            m.polygons[pumps[p]] = m.polygons[pumps[p]][::-1]


def variant_models(n_variants: int, min_facets: int = 2000, max_facets: int = 150000, seed: int = 42,
                   neg_range=(1e-4, 1e-1)):
    """Uneven beam-pipe variants: facet counts log-uniform in
    [min_facets, max_facets], NEG sticking log-uniform in neg_range."""
    rng = np.random.default_rng(seed)
    out = []
    for v in range(n_variants):
        target = math.exp(rng.uniform(math.log(min_facets), math.log(max_facets)))
        s = math.exp(rng.uniform(math.log(neg_range[0]), math.log(neg_range[1])))
        # main pipe carries ~85% of the facets: 2*n_az*n_ax triangles
        n_az = max(16, int(round(math.sqrt(0.85 * target / 2 / (340 / 128)))))
        n_ax = max(24, int(round(0.85 * target / 2 / n_az)))
        port_rings = max(2, int(round(16 * n_az / 128)))
        m = beam_pipe_model(n_az=n_az, n_ax=n_ax, port_rings=port_rings, neg_sticking=s,
                            n_textures=0, n_anglemaps=0, seed=seed + v)
        m.name = f"variant{v}_{m.n_facets}f_s{s:.2e}"
        out.append(m)
    return out


def adjoint_setup(m: Model, n_groups: int = 32, n_targets: int = 64, seed: int = 42,
                  s_range=(1e-3, 5e-2)):
    """Split the NEG facets into n_groups axial sub-bands with log-uniform
    base sticking, pick n_targets wall triangles outside the bands.
    Returns (group_of_facet, base_sticking, targets)."""
    rng = np.random.default_rng(seed)
    group = np.full(m.n_facets, -1, dtype=np.int32)
    neg = np.nonzero(m.neg_mask)[0]
    z = m.wall_z[neg]
    order = np.argsort(z, kind="stable")
    chunks = np.array_split(neg[order], n_groups)
    for g, ids in enumerate(chunks):
        group[ids] = g
    base = np.exp(rng.uniform(math.log(s_range[0]), math.log(s_range[1]), n_groups))
    walls = np.nonzero((m.kinds == 0) & (~m.neg_mask))[0]
    targets = np.sort(rng.choice(walls, n_targets, replace=False)).astype(np.int32)
    return group, base, targets
