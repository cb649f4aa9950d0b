"""mfg_intersect_batch_static, the fixed-ray-batch entry point that
allocates no device memory (rays and results staged through the library's
static device array), against mfg_intersect_batch on the same rays: the
results must be identical, on two calls in a row (the second smaller) and
with skip ids. Needs an MI355X."""
import numpy as np
import pytest

from helpers import random_rays
from molflow_sources_amd import _abi as A
from molflow_sources_amd import geometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    from molflow_sources_amd.simulation import GpuSimulation
    s = GpuSimulation(0)
    yield s
    s.close()


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("name", ["tube", "beam_pipe"])
def test_static_batch_equals_allocating_batch(sim, name):
    # origins inside the closed pipe (radius r_in), so that every ray hits a wall
    if name == "tube":
        m, lo, hi, r_in = geometry.tube_model(), [-0.9, -0.9, 0.2], [0.9, 0.9, 9.8], 0.9
    else:
        m, lo, hi, r_in = geometry.beam_pipe_model(n_textures=0, n_anglemaps=0), [-3, -3, 2], [3, 3, 198], 3.0
    sim.load(m)
    sim.rebuild_accel()

    def inside_rays(n, seed):
        o, d = random_rays(2 * n, lo, hi, seed=seed)
        keep = np.flatnonzero(o[:, 0] ** 2 + o[:, 1] ** 2 < r_in ** 2)[:n]
        assert len(keep) == n
        return np.ascontiguousarray(o[keep]), np.ascontiguousarray(d[keep])

    # first call at the maximum size, second one smaller, with skip ids
    o1, d1 = inside_rays(A.STATIC_MAX_RAYS, 3)
    r_alloc = sim.intersect(o1, d1)
    r_static = sim.intersect_static(o1, d1)
    assert np.all(r_alloc[0] >= 0)
    assert same(r_alloc, r_static)
    o2, d2 = inside_rays(5000, 4)
    skip = r_alloc[0][:5000].copy()
    r_alloc2 = sim.intersect(o2, d2, last_hit=skip)
    r_static2 = sim.intersect_static(o2, d2, last_hit=skip)
    assert same(r_alloc2, r_static2)
    # the second call's results are its own, not left over from the first
    assert not np.array_equal(r_static2[1], r_static[1][:5000])


def test_static_batch_refuses_more_than_its_maximum(sim):
    from molflow_sources_amd.simulation import MolflowGpuError
    sim.load(geometry.tube_model())
    sim.rebuild_accel()
    o, d = random_rays(A.STATIC_MAX_RAYS + 1, [-0.5, -0.5, 1], [0.5, 0.5, 9], seed=5)
    with pytest.raises(MolflowGpuError, match=r"\(6\)"):
        sim.intersect_static(o, d)
    # an empty batch is a no-op
    f, t, u, v = sim.intersect_static(np.zeros((0, 3)), np.zeros((0, 3)))
    assert len(f) == 0
