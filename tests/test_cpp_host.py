"""The C++ host shim (include/molflow_gpu.hpp) driven by tests/cpp/run_model:
the boundary used from C++ as a C++ simulation manager would use it."""
import os
import subprocess

import numpy as np
import pytest

from helpers import read_driver_output, write_model_file
from molflow_sources_amd import geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "run_model")


def test_driver_is_built_and_checks_its_arguments():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_model_file_round_trip_header(tmp_path):
    m = geometry.tube_model()
    p = tmp_path / "tube.bin"
    write_model_file(m, p)
    raw = p.read_bytes()
    assert raw[:8] == b"MFGM1\0\0\0"
    nv, nf = np.frombuffer(raw[8:16], dtype=np.uint32)
    assert nf == m.n_facets and nv == len(m.vertices)


@pytest.mark.gpu
@pytest.mark.parametrize("carry", [0, 1])
def test_cpp_driver_matches_python_runs(tmp_path, carry):
    from molflow_sources_amd.simulation import GpuSimulation
    m = geometry.tube_model()
    p, o = tmp_path / "tube.bin", tmp_path / f"out{carry}.txt"
    write_model_file(m, p)
    r = subprocess.run([DRIVER, str(p), str(o), "77", "3", "100000", str(carry)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    head, hits, absq = read_driver_output(o)
    s = GpuSimulation(0)
    s.load(m)
    s.rebuild_accel()
    s.reset(seed=77)
    s.run(300_000, max_events=1_000_000)
    t = s.read()
    s.close()
    assert head["desorbed"] == 300_000
    assert head["hits"] == int(t.glob[0]["nb_mc_hit"])
    assert np.array_equal(hits, t.facets["nb_mc_hit"][0].astype(np.int64))
    assert np.allclose(absq, t.facets["nb_abs_equiv"][0], rtol=0, atol=1e-9)
    if carry:
        assert head["in_flight_max"] > 0
