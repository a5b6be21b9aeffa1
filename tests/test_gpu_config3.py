"""The 96-wave-function batch, exactly as bench.py times it, against the oracle.

The batch (valence_amd.workloads.config3) is 48 displaced C8H18 geometries, 32
Givens-rotated C8H18 wave functions and 16 alkanes CnH2n+2 (n = 2..12), energy
and full gradient each, in one batch with graph replay.  References
(tests/golden/config3_fixtures.json, written by tests/golden/make_config3.py;
tests/golden/config_energies.json):
* energies of displaced members 0-3 and of all 16 alkanes: the oracle's, 1e-9 Eh;
* the 32 rotated members: config 1's oracle energy (a rotation of two doubly
  occupied orbitals leaves the energy unchanged), 1e-9 Eh;
* gradients: 32 weights of member 85 (C3H8) and directional derivatives of
  members 85, 48 (rotated C8H18) and 0 (displaced C8H18) against central
  differences of the oracle energy (h = 1e-5), 1e-7 norm-wise relative;
* every member: sum_k c_k dE/dc_k = 0 over each orbital's weights (the energy
  does not depend on an orbital's scale).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd import workloads

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(GOLDEN, "config3_fixtures.json")))
CONFIGS = json.load(open(os.path.join(GOLDEN, "config_energies.json")))


def direction(seed: int, n: int) -> np.ndarray:
    d = np.random.default_rng(seed).standard_normal(n)
    return d / np.linalg.norm(d)


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


@pytest.fixture(scope="module")
def batch_results(lib):
    systems = workloads.config3()
    assert len(systems) == 96
    b = lib.Batch(systems)
    b.set_profiling(True)
    print("config3: first (direct) run", flush=True)
    e0, g0 = b.run(want_grad=True)
    st0 = b.class_stats()
    assert b.uses_graph(), b.graph_error()
    print("config3: graph replays", flush=True)
    e1, g1 = b.run(want_grad=True)
    st1 = b.class_stats()
    e2, g2 = b.run(want_grad=True)
    st2 = b.class_stats()
    b.close()
    return systems, (e0, g0, st0), (e1, g1, st1), (e2, g2, st2)


def test_config3_replays_reproduce_direct_run(batch_results):
    systems, r0, r1, r2 = batch_results
    for e, g, _ in (r1, r2):
        assert np.all(np.isfinite(e))
        for i in range(len(systems)):
            assert abs(e[i] - r0[0][i]) < 1e-10, (i, e[i], r0[0][i])
            assert np.linalg.norm(g[i] - r0[1][i]) <= 1e-10 * np.linalg.norm(r0[1][i])


def test_config3_energies_vs_oracle(batch_results):
    systems, _, (e, _, _), _ = batch_results
    worst = 0.0
    for k, ent in FIX["energies"].items():
        i = int(k)
        assert systems[i].title == ent["title"]
        d = e[i] - ent["energy"]
        worst = max(worst, abs(d))
        print(f"  member {i:2d} {ent['title']:22s} gpu={e[i]:.10f} oracle={ent['energy']:.10f} diff={d:+.2e}")
        assert abs(d) < 1e-9, (i, e[i], ent["energy"])
    assert len(FIX["energies"]) >= 20
    e1 = CONFIGS["config1"]["energy"]
    for i in range(48, 80):
        assert systems[i].title.startswith("C8H18 rotated")
        worst = max(worst, abs(e[i] - e1))
        assert abs(e[i] - e1) < 1e-9, (i, e[i], e1)
    print(f"worst |E - oracle| over {len(FIX['energies']) + 32} members: {worst:.2e} Eh")


def test_config3_gradients_vs_oracle_fd(batch_results):
    systems, _, (_, g, _), _ = batch_results
    assert "85" in FIX["components"] and {"85", "48", "0"} <= set(FIX["directions"])
    for k, ent in FIX["components"].items():
        i = int(k)
        idx = np.array(sorted(int(x) for x in ent["fd"]))
        fd = np.array([ent["fd"][str(x)] for x in idx])
        rel = np.linalg.norm(g[i][idx] - fd) / np.linalg.norm(fd)
        print(f"  member {i} ({ent['title']}): {len(idx)} weights, rel err vs oracle FD {rel:.2e}")
        assert len(idx) >= 32 and rel < 1e-7
    for k, ent in FIX["directions"].items():
        i = int(k)
        gn = np.linalg.norm(g[i])
        for seed, fd in ent["fd"].items():
            gd = float(np.dot(g[i], direction(int(seed), systems[i].totlen)))
            err = abs(gd - fd) / gn
            print(f"  member {i} ({ent['title']}) direction {seed}: g.d={gd:.10e} fd={fd:.10e} err/|g|={err:.2e}")
            assert err < 1e-7


def test_config3_gradient_scale_identity(batch_results):
    systems, _, (_, g, _), _ = batch_results
    for i, s in enumerate(systems):
        for p in range(s.norbs):
            sl = slice(s.orb_start[p], s.orb_start[p + 1])
            v = float(np.dot(s.coeff[sl], g[i][sl]))
            assert abs(v) < 1e-8 * max(1.0, np.linalg.norm(g[i][sl]) * np.linalg.norm(s.coeff[sl])), (i, p, v)


def test_config3_replay_class_timing(batch_results):
    """Per-class device times of graph replays come from event nodes of the
    graph: finite, positive where the class ran, and not the direct run's."""
    _, (_, _, st0), (_, _, st1), (_, _, st2) = batch_results
    ran = st1[:, 5] > 0
    assert ran.sum() >= 15
    for st in (st1, st2):
        assert np.all(np.isfinite(st[:, 0]))
        assert np.all(st[ran, 0] > 0)
        assert np.all(st[ran, 1:5] == st0[ran, 1:5])   # counters: the same work every run
    assert not np.array_equal(st1[ran, 0], st0[ran, 0])
    assert not np.array_equal(st2[ran, 0], st1[ran, 0])
