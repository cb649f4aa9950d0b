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
    """The torch gradient of a scalar loss against finite differences of the
    same forward with common random numbers (same seed, same paths):
    Richardson-extrapolated central differences at h and 2h (h = 1e-2 s_g)
    remove the O(h^2) term; over 8 seeds the mean of the paired difference
    lies within 3.5 of its own standard errors. The gradient carries the
    low-flux cutoff's term; without it the differences show the cutoff's
    bias."""
    from molflow_sources_amd.adjoint import sticking_response
    sim, base, tg = unit
    n, nb = 10_000, 8
    w = torch.tensor(np.random.default_rng(4).normal(size=len(tg)))

    def loss(sv, seed, grad=False):
        s = torch.tensor(sv, dtype=torch.float64, requires_grad=grad)
        L = (sticking_response(s, sim, n, seed=seed, max_events=MAXEV) * w).sum()
        if grad:
            L.backward()
            return float(L.detach()), s.grad.detach().numpy().copy()
        return float(L.detach())

    for k in (0, len(base) // 2, len(base) - 1):
        h = 1e-2 * base[k]
        d = []
        for b in range(nb):
            seed = 100 + b
            _, g = loss(base, seed, grad=True)
            f = {}
            for q in (-2, -1, 1, 2):
                sv = base.copy()
                sv[k] += q * h
                f[q] = loss(sv, seed)
            d1 = (f[1] - f[-1]) / (2 * h)
            d2 = (f[2] - f[-2]) / (4 * h)
            d.append((4 * d1 - d2) / 3 - g[k])
        print(f"paired differences: group {k} done", flush=True)
        d = np.array(d)
        sem = d.std(ddof=1) / np.sqrt(nb)
        assert abs(d.mean()) <= 3.5 * sem + 1e-12 * np.abs(g).max(), (k, d.mean(), sem)


def test_backward_after_another_run_is_refused(unit):
    from molflow_sources_amd.adjoint import sticking_response
    from molflow_sources_amd.simulation import MolflowGpuError
    sim, base, tg = unit
    s = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    y = sticking_response(s, sim, 2000, seed=5, max_events=MAXEV)
    sim.reset(seed=1)
    with pytest.raises(MolflowGpuError):
        y.sum().backward()
