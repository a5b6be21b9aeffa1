"""torch.autograd binding (valence_amd.autograd) on the GPU: the device-pointer
entry against the host entry, backward against central differences of the
oracle energy (h = 1e-5, 1e-7 relative), and an L-BFGS run driven by the
adjoint gradient that must lower the energy monotonically."""
import json
import os

import numpy as np
import pytest
import torch

from valence_amd import workloads

pytestmark = pytest.mark.gpu

GRAD_FD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grad_fd.json")))


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


@pytest.mark.parametrize("which", [0, 2])
def test_device_entry_matches_host(lib, which):
    from valence_amd.autograd import energy
    s = workloads.CONFIGS[which]()
    with lib.Context(s) as ctx:
        e_h, g_h = ctx.energy_grad()
        c = torch.tensor(s.coeff, dtype=torch.float64, device="cuda", requires_grad=True)
        e = energy(ctx, c)
        e.backward()
        torch.cuda.synchronize()
    print(f"config{which}: host {e_h:.12f} device {float(e):.12f}")
    assert abs(float(e) - e_h) < 1e-10
    assert torch.linalg.norm(c.grad.cpu() - torch.from_numpy(g_h)) <= 1e-10 * np.linalg.norm(g_h)


def test_backward_vs_finite_differences(lib):
    """backward() against central differences of the oracle energy (h = 1e-5,
    every weight of the water workload, tests/golden/grad_fd.json)."""
    from valence_amd.autograd import energy
    s = workloads.config0()
    d = GRAD_FD["config0"]["fd"]
    idx = np.array(sorted(int(k) for k in d))
    fd = np.array([d[str(k)] for k in idx])
    assert len(idx) == s.totlen
    with lib.Context(s) as ctx:
        c = torch.tensor(s.coeff, dtype=torch.float64, device="cuda", requires_grad=True)
        (3.0 * energy(ctx, c)).backward()          # chain rule through a scale
        g = c.grad.cpu().numpy() / 3.0
    rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
    print(f"rel err vs oracle FD = {rel:.3e}")
    assert rel < 1e-7


def test_lbfgs_lowers_the_energy(lib):
    from valence_amd.autograd import optimize_weights
    s = workloads.config0()
    hist = []
    c = optimize_weights(s, steps=8, device="cuda", history=hist)
    with lib.Context(s) as ctx:
        e0 = ctx.energy()
        e1 = ctx.energy(c)
    print(f"E0={e0:.10f} after={e1:.10f} history={['%.6f' % h for h in hist]}")
    assert e1 < e0 - 1e-3
    assert all(b <= a + 1e-10 for a, b in zip(hist, hist[1:]))
