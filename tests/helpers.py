"""Shared helpers of the parity tests."""
from __future__ import annotations

import math

import numpy as np


def random_rays(n, lo, hi, seed=1):
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def canonical_trace(facets, kinds):
    """Sort runs of consecutive transparent passes (kind 3) by facet id: the
    BVH walk order of passes between two hard events is not part of the
    physics."""
    f = facets.copy()
    k = kinds.copy()
    for row in range(f.shape[0]):
        j = 0
        n = f.shape[1]
        while j < n:
            if k[row, j] == 3:
                e = j
                while e < n and k[row, e] == 3:
                    e += 1
                f[row, j:e] = np.sort(f[row, j:e])
                j = e
            else:
                j += 1
    return f, k


def z_two_proportions(k1, n1, k2, n2):
    p = (k1 + k2) / (n1 + n2)
    s = math.sqrt(max(p * (1 - p) * (1 / n1 + 1 / n2), 1e-300))
    return (k1 / n1 - k2 / n2) / s


def z_means(m1, s1, m2, s2):
    return (m1 - m2) / math.sqrt(max(s1 ** 2 + s2 ** 2, 1e-300))


def write_model_file(model, path):
    """Model file read by tests/cpp/run_model (single structure, no
    parameters or moments): header, vertices, polygon offsets and indices,
    then the raw facet and settings structs of molflow_gpu.h."""
    import ctypes as C
    import struct

    import numpy as np
    v = model.view()
    nv, nf = v.n_vertices, v.n_facets
    off = np.ctypeslib.as_array(v.facet_vertex_offset, shape=(nf + 1,)).astype(np.uint32)
    idx = np.ctypeslib.as_array(v.facet_vertex_index, shape=(int(off[-1]),)).astype(np.uint32)
    verts = np.ctypeslib.as_array(v.vertices, shape=(3 * nv,)).astype(np.float64)
    from molflow_sources_amd._abi import FacetDesc, Settings
    fbytes = C.string_at(v.facets, nf * C.sizeof(FacetDesc))
    sbytes = C.string_at(v.settings, C.sizeof(Settings))
    with open(path, "wb") as fh:
        fh.write(b"MFGM1\0\0\0")
        fh.write(struct.pack("<4I", nv, nf, len(idx), 0))
        fh.write(verts.tobytes())
        fh.write(off.tobytes())
        fh.write(idx.tobytes())
        fh.write(fbytes)
        fh.write(sbytes)


def read_driver_output(path):
    import numpy as np
    head, rows = {}, []
    for line in open(path):
        parts = line.split()
        if parts[0].isdigit():
            rows.append((int(parts[1]), float(parts[2])))
        else:
            head[parts[0]] = float(parts[1]) if "." in parts[1] or "e" in parts[1] else int(parts[1])
    hits = np.array([r[0] for r in rows], dtype=np.int64)
    absq = np.array([r[1] for r in rows])
    return head, hits, absq


def polygon_zoo_model():
    """Four large planar polygons stacked along z (2-sided, sticking 0.3):
    an 11-pointed star (non-convex, 22 vertices, the desorbing source), an
    axis-aligned comb (equal vertex heights and horizontal edges), a 64-gon
    and a C shape. They exercise the point-in-polygon test of polygons with
    more than 4 vertices, in intersections and in desorption sampling."""
    from molflow_sources_amd._abi import DES_COSINE
    from molflow_sources_amd.model import Model, default_facets

    k = 11
    star = [((1.0 if i % 2 == 0 else 0.45) * math.cos(math.pi * i / k),
             (1.0 if i % 2 == 0 else 0.45) * math.sin(math.pi * i / k)) for i in range(2 * k)]
    comb = [(0, 0), (4, 0), (4, 1), (3.5, 1), (3.5, 0.4), (2.5, 0.4), (2.5, 1), (1.5, 1), (1.5, 0.4),
            (0.5, 0.4), (0.5, 1), (0, 1)]
    comb = [(0.5 * x - 1.0, 1.5 * y - 0.75) for x, y in comb]
    disk = [(math.cos(2 * math.pi * i / 64), math.sin(2 * math.pi * i / 64)) for i in range(64)]
    c_shape = ([(math.cos(a), math.sin(a)) for a in np.linspace(0.4, 2 * math.pi - 0.4, 10)] +
               [(0.5 * math.cos(a), 0.5 * math.sin(a)) for a in np.linspace(2 * math.pi - 0.4, 0.4, 10)])
    verts, polys = [], []
    for z, shape in enumerate((star, comb, disk, c_shape)):
        base = len(verts)
        verts += [(x, y, float(z)) for x, y in shape]
        polys.append(list(range(base, base + len(shape))))
    F = default_facets(len(polys))
    F["is_2sided"] = 1
    F["sticking"] = 0.3
    F["outgassing"][0] = 1.0
    F["desorb_type"][0] = DES_COSINE
    return Model(np.array(verts, dtype=float), polys, F, name="polygon_zoo")
