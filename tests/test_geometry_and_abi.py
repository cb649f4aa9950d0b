"""Synthetic workloads, the C ABI surface, and host logic (no GPU)."""
import ctypes as C
import os
import re
import subprocess
from collections import Counter

import numpy as np
import pytest

from molflow_sources_amd import _abi as A
from molflow_sources_amd import geometry
from molflow_sources_amd.model import parse_user_moment
from oracle.mfcpu import CpuModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def edges_shared_twice(m):
    cnt = Counter()
    for p in m.polygons:
        for i in range(len(p)):
            a, b = p[i], p[(i + 1) % len(p)]
            cnt[(min(a, b), max(a, b))] += 1
    return all(v == 2 for v in cnt.values()), cnt


def test_tube_model_layout(oracle_lib):
    m = geometry.tube_model()
    assert m.n_facets == 3202
    ok, _ = edges_shared_twice(m)
    assert ok
    cm = CpuModel(m)
    # inward normals: wall facet normals point to the axis, inlet +z, outlet -z
    assert cm.frame(0)[3][2] == pytest.approx(1.0)
    assert cm.frame(1)[3][2] == pytest.approx(-1.0)
    fr = cm.frame(2)
    centre = fr[0] + 0.5 * fr[1] + 0.5 * fr[2]
    assert np.dot(fr[3], -np.array([centre[0], centre[1], 0])) > 0
    # 64-gon area and wall area
    assert cm.area(0) == pytest.approx(0.5 * 64 * np.sin(2 * np.pi / 64), rel=1e-12)
    assert sum(cm.area(f) for f in range(2, 3202)) == pytest.approx(64 * 2 * np.sin(np.pi / 64) * 10, rel=1e-9)


def test_beam_pipe_is_watertight_and_sized(oracle_lib):
    m = geometry.beam_pipe_model()
    assert 95_000 < m.n_facets < 110_000
    ok, cnt = edges_shared_twice(m)
    assert ok, [e for e, v in cnt.items() if v != 2][:5]
    assert (m.facets["is_textured"] == 1).sum() == 16
    assert (m.facets["am_record"] == 1).sum() == 4
    assert np.all(m.facets["sticking"][m.pump_ids] == 0.05)
    assert abs(m.neg_mask.sum() * 2 * np.pi * 4 / (m.neg_mask.sum()) - 2 * np.pi * 4) < 1e-9
    # a closed model: no ray from inside may leak (oracle brute force)
    cm = CpuModel(m)
    rng = np.random.default_rng(3)
    n = 2000
    z = rng.uniform(1, 199, n)
    r = rng.uniform(0, 3.5, n)
    a = rng.uniform(0, 2 * np.pi, n)
    o = np.stack([r * np.cos(a), r * np.sin(a), z], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f, t, _, _ = cm.intersect(o, d, brute=False)
    assert np.all(f >= 0)


def test_time_dependent_tube(oracle_lib):
    m = geometry.time_dependent_tube()
    assert len(m.moments) == 100
    assert m.moments[0][0] == pytest.approx(0.25e-3) and m.moments[-1][0] == pytest.approx(49.75e-3, rel=1e-9)
    cm = CpuModel(m)
    assert cm.latest_moment == pytest.approx(50e-3, rel=1e-9)
    assert cm.layout().n_slots == 101


def test_variants_are_uneven(oracle_lib):
    vs = geometry.variant_models(6, min_facets=2000, max_facets=20000, seed=1)
    sizes = [v.n_facets for v in vs]
    assert min(sizes) >= 1500 and max(sizes) <= 30000 and len(set(sizes)) == 6


def test_adjoint_setup_groups(oracle_lib):
    m = geometry.beam_pipe_model(n_az=32, n_ax=80, port_rings=4, n_textures=0, n_anglemaps=0)
    g, s, t = geometry.adjoint_setup(m, n_groups=32, n_targets=64)
    assert g.max() == 31 and len(set(g[g >= 0])) == 32
    assert np.all((s >= 1e-3) & (s <= 5e-2))
    assert len(set(t)) == 64 and np.all(g[t] == -1)


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "molflow_gpu.h")).read()
    return sorted(set(re.findall(r"\b(mfg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "molflow_sources_amd", "libmolflow_gpu.so")
    assert os.path.exists(lib), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (mfg_\w+)", out))
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    assert set(A.EXPORTED_SYMBOLS) == set(decl)


def test_library_loads_and_reports_abi():
    lib = A.load_library()
    assert lib.mfg_abi_version() == A.MFG_ABI_VERSION


def test_struct_sizes_match_the_c_header(oracle_lib):
    # the oracle (plain C++) rejects a view whose struct_size differs from
    # sizeof(mfg_model_view); loading succeeds only if the ctypes mirror matches
    m = geometry.tube_model(n_az=8, n_ax=2)
    CpuModel(m)
    assert C.sizeof(A.FacetDesc) % 8 == 0


def test_python_moment_parser_matches_oracle(oracle_lib):
    from oracle import mfcpu
    for c, w in (("1e-3,1e-3,4e-2", 1e-3), ("0.25e-3,0.5e-3,49.8e-3", 5e-4), ("0", 1.0)):
        assert np.array_equal(np.array(parse_user_moment(c, w)), mfcpu.parse_user_moment(c, w))


def test_microbench_library_exports_its_header():
    """The measurement helper (atomic throughput, random-line read rate)
    loads and exports the symbols its header declares; no GPU call is made."""
    import ctypes
    import re
    lib_path = os.path.join(ROOT, "molflow_sources_amd", "libmfg_microbench.so")
    if not os.path.exists(lib_path):
        pytest.skip("libmfg_microbench.so not built")
    txt = open(os.path.join(ROOT, "include", "mfg_microbench.h")).read()
    names = re.findall(r"^int\s+(mfgb_\w+)\(", txt, flags=re.M)
    assert names == ["mfgb_atomic_rates", "mfgb_gather_rate"]
    lib = ctypes.CDLL(lib_path)
    for n in names:
        assert hasattr(lib, n)


def _bad_models():
    """(name, model) pairs that the loader must refuse."""
    out = []
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["super_idx"][5] = -2          # below the "every structure" value
    out.append(("super_idx -2", m))
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["super_idx"][0] = 1           # the only structure is 0
    out.append(("super_idx past the structures", m))
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["teleport_dest"][3] = -3      # -1 is the lowest meaning
    out.append(("teleport_dest -3", m))
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["teleport_dest"][3] = m.n_facets + 1
    out.append(("teleport_dest past the facets", m))
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["super_dest"][3] = 2          # link to a structure that does not exist
    out.append(("super_dest past the structures", m))
    return out


@pytest.mark.parametrize("name,model", _bad_models(), ids=[n for n, _ in _bad_models()])
def test_out_of_range_indices_are_refused_at_load(oracle_lib, name, model):
    """The library's load-time preparation (through mfg_sanity_check, which
    runs it without a device) and the oracle both refuse facet indices that
    the kernel would otherwise use to index device tables."""
    from molflow_sources_amd.simulation import sanity_check
    problems = sanity_check(model)
    assert problems, name
    with pytest.raises(ValueError):
        CpuModel(model)


def test_sanity_check_accepts_good_models_and_reports_reference_checks(oracle_lib):
    from molflow_sources_amd.simulation import sanity_check
    assert sanity_check(geometry.tube_model(n_az=8, n_ax=2)) == []
    assert sanity_check(geometry.time_dependent_tube()) == []
    # MolflowSimulationModel::SanityCheck (MolflowSimGeom.cpp:337-404) items
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["count_abs"][1] = 1           # texture counter on an untextured facet
    assert any("Untextured facet with texture counters" in p for p in sanity_check(m))
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.facets["outgassing"][0] = 0.0
    m.facets["desorb_type"][0] = 0
    probs = sanity_check(m)
    assert any("No desorption facet found" in p for p in probs)
    assert any("desorbing facet" in p for p in probs)   # the load itself fails too
    m = geometry.tube_model(n_az=8, n_ax=2)
    m.settings.enable_decay = 1
    m.settings.half_life = 0.0
    assert any("half life" in p for p in sanity_check(m))


def test_oracle_adjoint_refuses_sticking_outside_the_open_interval(oracle_lib):
    """d/ds = -n/(1-s) + 1/s is undefined at s = 0 and s = 1: the oracle's
    adjoint refuses such group values instead of returning NaN or dropping
    the -n/(1-s) term."""
    import ctypes as C2
    from oracle import mfcpu
    m = geometry.beam_pipe_model(n_az=16, n_ax=20, port_rings=2, n_textures=0, n_anglemaps=0)
    m.settings.low_flux_mode = 1
    g, s, t = geometry.adjoint_setup(m, n_groups=4, n_targets=8)
    cm = CpuModel(m)
    L = mfcpu.lib()
    j0, jac = np.zeros(8), np.zeros(32)
    gi, ti = np.ascontiguousarray(g, np.int32), np.ascontiguousarray(t, np.int32)
    for bad in (0.0, 1.0):
        sv = np.array([0.01, bad, 0.02, 0.03])
        rc = L.mfcpu_adjoint_run(cm.h, 42, 0, 10, gi.ctypes.data_as(C2.POINTER(C2.c_int32)), 4,
                                 sv.ctypes.data_as(C2.POINTER(C2.c_double)),
                                 ti.ctypes.data_as(C2.POINTER(C2.c_int32)), 8, 1,
                                 j0.ctypes.data_as(C2.POINTER(C2.c_double)),
                                 jac.ctypes.data_as(C2.POINTER(C2.c_double)), 1000)
        assert rc != 0, bad
    sv = np.array([0.01, 0.5, 0.02, 0.03])
    rc = L.mfcpu_adjoint_run(cm.h, 42, 0, 10, gi.ctypes.data_as(C2.POINTER(C2.c_int32)), 4,
                             sv.ctypes.data_as(C2.POINTER(C2.c_double)), ti.ctypes.data_as(C2.POINTER(C2.c_int32)), 8, 1,
                             j0.ctypes.data_as(C2.POINTER(C2.c_double)), jac.ctypes.data_as(C2.POINTER(C2.c_double)),
                             1000)
    assert rc == 0 and np.all(np.isfinite(jac))


def test_kernel_hash_ignores_comments_only():
    """bench.py keys the profile summaries to the code of the kernel
    sources: comments and blank lines do not change the key, code does, and
    comment markers inside literals are code."""
    import bench
    a = 'int a = 1; // note\n/* block\n comment */ char q = \'/\';\n\nconst char* t = "x // y";\n'
    b = 'int a = 1;\n char q = \'/\';\nconst char* t = "x // y";   \n'
    assert bench.code_only(a) == bench.code_only(b)
    assert bench.code_only(a) != bench.code_only(a.replace("a = 1", "a = 2"))
    assert "x // y" in bench.code_only(a)
