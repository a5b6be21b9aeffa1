"""GPU parity of the spin-coupled (npair > 0) energy and adjoint gradient, through the C ABI.

Every recorded spin-coupled case of the reference (testing/testing.py:69-150,
examples/test_examples.py:63-152) against the oracle (1e-9 Eh) and the recorded
energy (1e-8 Eh); the 12-atom H chain with six Rumer pairs (64 determinants)
against its oracle energy (1e-9 Eh); gradients against central differences of
the oracle energy (h = 1e-5, 1e-7 norm-wise relative).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd.system import read_input
from valence_amd import workloads

pytestmark = pytest.mark.gpu

KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))
ORACLE_KNOWN = json.load(open(os.path.join(GOLDEN, "oracle_known_answers.json")))
CONFIGS = json.load(open(os.path.join(GOLDEN, "config_energies.json")))
GRAD_FD = json.load(open(os.path.join(GOLDEN, "grad_fd.json")))


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


def _load(key):
    grp, name = key.split("/")
    return read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{name}"))


SC = sorted(k for k in ORACLE_KNOWN if _load(k).npair > 0)
# testing/n2.sc4val-b.p2 has determinant pairs whose transition overlaps are
# exactly singular (orthogonal orbitals by symmetry; corank 1 and 2).


@pytest.mark.parametrize("key", SC)
def test_sc_known_answers_gpu(lib, key):
    grp, name = key.split("/")
    print(f"case {key}", flush=True)
    s = _load(key)
    with lib.Context(s) as ctx:
        e = ctx.energy()
    ref = KNOWN[grp][name]["guess_energy"]
    orc = ORACLE_KNOWN[key]["oracle_energy"]
    print(f"  gpu={e:.16f} known={ref:.16f} oracle={orc:.16f} d_known={e - ref:+.3e} d_oracle={e - orc:+.3e}")
    assert abs(e - orc) < 1e-9
    assert abs(e - ref) < 1e-8


def test_config2_energy(lib):
    """H12 chain, six Rumer pairs: 64 x 64 determinant pairs."""
    s = workloads.config2()
    g = CONFIGS["config2"]
    with lib.Context(s) as ctx:
        e = ctx.energy()
        e2 = ctx.energy()
        eg, _ = ctx.energy_grad()
    print(f"gpu={e:.12f} oracle={g['energy']:.12f} unscreened={g['energy_unscreened']:.12f}")
    assert abs(e - e2) < 1e-12  # Coulomb accumulation uses atomics: summation order varies
    assert abs(eg - e) < 1e-10
    assert abs(e - g["energy_unscreened"]) < 1e-9
    assert abs(e - g["energy"]) < 1e-9


@pytest.mark.parametrize("which", ["config2", "testing/be-sc", "testing/h2o-vdz-sc1", "examples/lih.SCval"])
def test_sc_gradient_fd(lib, which):
    """The adjoint gradient against central differences of the oracle energy
    (h = 1e-5, tests/golden/grad_fd.json): 24 sampled weights of config 2, every
    weight of the recorded inputs."""
    print(f"case {which}", flush=True)
    s = workloads.config2() if which == "config2" else _load(which)
    if which != "config2":
        # the recorded inputs are optimized (|dE/dc| ~ 1e-5, below what central
        # differences resolve): test at a seeded perturbation of the weights
        rng = np.random.default_rng(7)
        s.coeff = s.coeff + 0.05 * rng.standard_normal(s.totlen) * np.maximum(np.abs(s.coeff), 0.1)
    d = GRAD_FD[which]["fd"]
    idx = np.array(sorted(int(k) for k in d))
    fd = np.array([d[str(k)] for k in idx])
    assert len(idx) == (24 if which == "config2" else s.totlen)
    with lib.Context(s) as ctx:
        e, g = ctx.energy_grad()
    rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
    print(f"  E={e:.12f} |g|={np.linalg.norm(g):.6e} rel err vs oracle FD={rel:.3e}")
    assert rel < 1e-7


def test_sc_gradient_scale_invariance(lib):
    """E is invariant to scaling one orbital: sum_k c_k dE/dc_k = 0 over each orbital's weights."""
    s = workloads.config2()
    with lib.Context(s) as ctx:
        _, g = ctx.energy_grad()
    for p in range(s.norbs):
        sl = slice(s.orb_start[p], s.orb_start[p + 1])
        v = float(np.dot(s.coeff[sl], g[sl]))
        assert abs(v) < 1e-9 * max(1.0, np.linalg.norm(g[sl]) * np.linalg.norm(s.coeff[sl]))


def test_sc_batch_matches_single(lib):
    s2 = workloads.config2()
    s0 = workloads.config0()
    b = lib.Batch([s2, s0])
    e, g = b.run(want_grad=True)
    for k, s in enumerate([s2, s0]):
        with lib.Context(s) as ctx:
            e1, g1 = ctx.energy_grad()
        assert abs(e[k] - e1) < 1e-10
        assert np.linalg.norm(g[k] - g1) <= 1e-10 * np.linalg.norm(g1)
