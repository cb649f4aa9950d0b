"""torch.autograd binding of the sticking adjoint (molflow_sources_amd.adjoint).
Needs an MI355X."""
import numpy as np
import pytest
import torch

from molflow_sources_amd import geometry

pytestmark = pytest.mark.gpu

MAXEV = 5_000_000


def small_adjoint_model():
    m = geometry.beam_pipe_model(n_az=48, n_ax=120, port_rings=6, n_textures=0, n_anglemaps=0)
    m.settings.low_flux_mode = 1
    m.settings.low_flux_cutoff = 1e-6
    g, s, t = geometry.adjoint_setup(m, n_groups=8, n_targets=16, seed=5)
    return m, g, s, t


@pytest.fixture(scope="module")
def unit():
    from molflow_sources_amd.simulation import GpuSimulation
    m, g, s, t = small_adjoint_model()
    sim = GpuSimulation(0)
    sim.load(m)
    sim.rebuild_accel()
    sim.adjoint_setup(g, t)
    yield sim, s, t
    sim.close()


def test_autograd_gradient_equals_path_score_contraction(unit):
    from molflow_sources_amd.adjoint import sticking_response
    sim, base, tg = unit
    n = 20_000
    w = torch.tensor(np.random.default_rng(3).normal(size=len(tg)))
    scale = torch.linspace(0.5, 2.0, len(tg), dtype=torch.float64)
    s = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    y = sticking_response(s, sim, n, seed=17, max_events=MAXEV, scale=scale)
    (y * w).sum().backward()
    j0, jac = sim.adjoint_read()
    assert np.allclose(y.detach().numpy(), scale.numpy() * j0 / n, rtol=1e-12, atol=0)
    expect = ((w * scale).numpy() @ jac) / n
    assert np.allclose(s.grad.numpy(), expect, rtol=1e-10, atol=1e-15 * np.abs(expect).max())
    assert np.all(np.isfinite(s.grad.numpy())) and np.abs(s.grad.numpy()).max() > 0


def test_autograd_matches_finite_difference_of_its_forward(unit):
    """The torch gradient of a scalar loss against central differences of
    the same forward with common random numbers (same seed, same paths)."""
    from molflow_sources_amd.adjoint import sticking_response
    sim, base, tg = unit
    n = 20_000
    w = torch.tensor(np.random.default_rng(4).normal(size=len(tg)))

    def loss(sv, grad=False):
        s = torch.tensor(sv, dtype=torch.float64, requires_grad=grad)
        L = (sticking_response(s, sim, n, seed=23, max_events=MAXEV) * w).sum()
        if grad:
            L.backward()
            return float(L.detach()), s.grad.detach().numpy().copy()
        return float(L.detach())

    _, g = loss(base, grad=True)
    for k in (0, len(base) // 2, len(base) - 1):
        ds = 1e-3 * base[k]
        sp, sm = base.copy(), base.copy()
        sp[k] += ds
        sm[k] -= ds
        fd = (loss(sp) - loss(sm)) / (2 * ds)
        # common random numbers: the paths are shared, only weights and the
        # low-flux cutoff point move; that cutoff leaves an absolute error
        # of order 1e-5 of the gradient's scale
        assert abs(fd - g[k]) <= 2e-3 * abs(g[k]) + 1e-5 * np.abs(g).max(), (k, fd, g[k])


def test_backward_after_another_run_is_refused(unit):
    from molflow_sources_amd.adjoint import sticking_response
    from molflow_sources_amd.simulation import MolflowGpuError
    sim, base, tg = unit
    s = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    y = sticking_response(s, sim, 2000, seed=5, max_events=MAXEV)
    sim.reset(seed=1)
    with pytest.raises(MolflowGpuError):
        y.sum().backward()
