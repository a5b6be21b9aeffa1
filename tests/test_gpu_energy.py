"""GPU parity of the plain energy and the adjoint gradient (HIP path through the C ABI).

References: the oracle (C restatement of the reference algorithm, pinned by the
reference's 58 recorded energies) and its committed values under tests/golden/.
Tolerances: energies 1e-9 Eh against the oracle run without the reference's
orbital-quartet screens (the HIP path evaluates every quartet above a 1e-17
AO Schwarz bound) and with them; gradients 1e-7 norm-wise relative against
central finite differences of the oracle energy with h = 1e-5
(tests/golden/grad_fd.json).
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


def _closed_or_high_spin(key):
    grp, name = key.split("/")
    s = read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{name}"))
    return s if s.npair == 0 else None


NONSC = sorted(k for k in ORACLE_KNOWN if _closed_or_high_spin(k) is not None)


@pytest.mark.parametrize("key", NONSC)
def test_known_answers_gpu(lib, key):
    grp, name = key.split("/")
    print(f"case {key}", flush=True)
    s = read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{name}"))
    with lib.Context(s) as ctx:
        e = ctx.energy()
    ref = KNOWN[grp][name]["guess_energy"]
    orc = ORACLE_KNOWN[key]["oracle_energy"]
    print(f"  gpu={e:.16f} known={ref:.16f} oracle={orc:.16f} d_known={e - ref:+.3e} d_oracle={e - orc:+.3e}")
    assert abs(e - ref) < 1e-8
    assert abs(e - orc) < 1e-9


def test_config0_energy(lib):
    s = workloads.config0()
    g = CONFIGS["config0"]
    with lib.Context(s) as ctx:
        e = ctx.energy()
        e2 = ctx.energy()
    print(f"gpu={e:.12f} oracle={g['energy']:.12f} unscreened={g['energy_unscreened']:.12f}")
    assert e == e2  # deterministic reduction
    assert abs(e - g["energy_unscreened"]) < 1e-9
    assert abs(e - g["energy"]) < 1e-9


def _unique_quartets(s):
    """T(T+1)/2 over the T = nsh(nsh+1)/2 shell pairs: each unique shell quartet
    is evaluated or skipped by the Schwarz cut exactly once."""
    nsh = sum(int(s.type_shell_start[t + 1] - s.type_shell_start[t]) for t in s.atom_type)
    T = nsh * (nsh + 1) // 2
    return T * (T + 1) // 2


def oracle_fd(name):
    """Central differences of the oracle energy (h = 1e-5, tests/golden/make_grad_fd.py):
    weight indices and dE/dc values."""
    d = GRAD_FD[name]["fd"]
    idx = np.array(sorted(int(k) for k in d))
    return idx, np.array([d[str(k)] for k in idx])


def test_config0_gradient_fd(lib):
    """Every weight's derivative against central differences of the oracle energy."""
    s = workloads.config0()
    with lib.Context(s) as ctx:
        e, g = ctx.energy_grad()
        e0 = ctx.energy(s.coeff)
    idx, fd = oracle_fd("config0")
    assert len(idx) == s.totlen
    print(f"E(grad call)={e:.12f} E(energy call)={e0:.12f}")
    assert abs(e - e0) < 1e-9
    rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
    print(f"|g|={np.linalg.norm(g):.6e} rel err vs oracle FD={rel:.3e}")
    assert rel < 1e-7


def test_gradient_scale_invariance(lib):
    """E is invariant to scaling each orbital, so sum_k c_k dE/dc_k = 0 per orbital."""
    s = workloads.config0()
    with lib.Context(s) as ctx:
        _, g = ctx.energy_grad()
    for p in range(s.norbs):
        sl = slice(s.orb_start[p], s.orb_start[p + 1])
        v = float(np.dot(s.coeff[sl], g[sl]))
        assert abs(v) < 1e-8 * max(1.0, np.linalg.norm(g[sl]) * np.linalg.norm(s.coeff[sl]))


def test_config1_energy(lib):
    """C8H18 (210 AOs, 33 docc) against the oracle's full energy."""
    s = workloads.config1()
    g = CONFIGS["config1"]
    with lib.Context(s) as ctx:
        e = ctx.energy()
        st = ctx.counters()
    print(f"gpu={e:.12f} oracle={g['energy']:.12f} unscreened={g['energy_unscreened']:.12f} "
          f"quartets={st[0]:.0f} skipped={st[2]:.0f} eri_ms={st[3]:.2f}")
    assert abs(e - g["energy_unscreened"]) < 1e-9
    assert abs(e - g["energy"]) < 1e-9


def test_grid_inverse_config1(lib):
    """The grid-wide Gauss-Jordan inverse (one pair of launches per pivot), which
    the default runs only for overlap blocks above 128 rows, forced on C8H18's
    33 x 33 block: the oracle's energy to 1e-9 Eh, the one-workgroup inverse's
    energy and gradient to rounding, and P S_orb = 1 at FP64 level."""
    s = workloads.config1()
    g = CONFIGS["config1"]
    with lib.Context(s) as ctx:
        e0, g0 = ctx.energy_grad()
    with lib.Context(s, grid_inverse_min=1) as ctx:
        e = ctx.energy()
        e1, g1 = ctx.energy_grad()
        n = s.nspatial
        S = ctx.buffer("S_orb", n * n).reshape(n, n)
        P = ctx.buffer("P", n * n).reshape(n, n)
    res = np.abs(P @ S - np.eye(n)).max()
    print(f"grid inverse: E={e:.12f} oracle={g['energy']:.12f} one-workgroup={e0:.12f} |PS-1|max={res:.2e}")
    assert abs(e - g["energy"]) < 1e-9 and abs(e - g["energy_unscreened"]) < 1e-9
    assert abs(e1 - e0) < 1e-10
    assert np.linalg.norm(g1 - g0) <= 1e-10 * np.linalg.norm(g0)
    assert res < 1e-12
    with pytest.raises(lib.VsvbError, match="grid_inverse_min"):
        lib.Context(s, grid_inverse_min=-1)


def test_rotation_invariance(lib):
    """Mixing doubly occupied orbitals leaves a VSHF energy unchanged (size-independent check)."""
    s = workloads.config1()
    with lib.Context(s) as ctx:
        e0 = ctx.energy()
    r = workloads.rotate_pair(s, 3, 17, 0.04)
    r = workloads.rotate_pair(r, 0, 30, -0.03)
    with lib.Context(r) as ctx:
        e1 = ctx.energy()
    assert abs(e1 - e0) < 1e-9 * max(1.0, abs(e0) / 1e4)


def test_batch_matches_single(lib):
    systems = workloads.config3()
    pick = [0, 50, 90, 95]
    b = lib.Batch([systems[i] for i in pick])
    e, g = b.run(want_grad=True)
    for k, i in enumerate(pick):
        with lib.Context(systems[i]) as ctx:
            e1, g1 = ctx.energy_grad()
            e2 = ctx.energy()
        assert abs(e[k] - e1) < 1e-9 and abs(e[k] - e2) < 1e-9
        assert np.linalg.norm(g[k] - g1) <= 1e-10 * np.linalg.norm(g1)


def test_batch_groups_share_integrals(lib):
    """Wave functions on one molecule share a single ERI pass; results must equal
    independent evaluations (rotations of C8H18 and one displaced geometry)."""
    systems = workloads.config3()
    pick = [48, 49, 50, 60, 0, 90]          # four rotations (one group), one displaced, one alkane
    b = lib.Batch([systems[i] for i in pick])
    e, g = b.run(want_grad=True)
    for k, i in enumerate(pick):
        with lib.Context(systems[i]) as ctx:
            e1, g1 = ctx.energy_grad()
        print(f"wfn {i}: batch {e[k]:.12f} single {e1:.12f}")
        assert abs(e[k] - e1) < 1e-9
        assert np.linalg.norm(g[k] - g1) <= 1e-10 * np.linalg.norm(g1)


def test_batch_graph_replays(lib):
    """The first batch run is direct and then captured into a HIP graph; replays,
    with direct single evaluations between them, must reproduce it."""
    systems = workloads.config3()
    pick = [0, 48, 49, 90]
    b = lib.Batch([systems[i] for i in pick])
    e0, g0 = b.run(want_grad=True)
    assert b.uses_graph(), b.graph_error()
    for rep in range(3):
        with lib.Context(systems[pick[rep]]) as ctx:
            ctx.energy_grad()                      # direct work between replays
        e, g = b.run(want_grad=True)
        for k in range(len(pick)):
            assert abs(e[k] - e0[k]) < 1e-10, (rep, k, e[k], e0[k])
            assert np.linalg.norm(g[k] - g0[k]) <= 1e-10 * np.linalg.norm(g0[k])
    for k, i in enumerate(pick):
        with lib.Context(systems[i]) as ctx:
            e1, g1 = ctx.energy_grad()
        assert abs(e[k] - e1) < 1e-9
        assert np.linalg.norm(g[k] - g1) <= 1e-10 * np.linalg.norm(g1)


def test_batch_graph_off_matches(lib):
    """With graph replay switched off (before the first run, and again after a
    graph exists) every run is direct and reproduces the replayed results."""
    systems = workloads.config3()
    pick = [1, 50, 92]
    b = lib.Batch([systems[i] for i in pick])
    e0, g0 = b.run(want_grad=True)
    e1, g1 = b.run(want_grad=True)
    assert b.uses_graph(), b.graph_error()
    b.set_graph(False)
    assert not b.uses_graph()
    for rep in range(2):
        e, g = b.run(want_grad=True)
        assert not b.uses_graph()
        for k in range(len(pick)):
            assert abs(e[k] - e1[k]) < 1e-10, (rep, k, e[k], e1[k])
            assert np.linalg.norm(g[k] - g1[k]) <= 1e-10 * np.linalg.norm(g1[k])
    b2 = lib.Batch([systems[i] for i in pick])
    b2.set_graph(False)
    for rep in range(2):
        e, g = b2.run(want_grad=True)
        assert not b2.uses_graph()
        for k in range(len(pick)):
            assert abs(e[k] - e0[k]) < 1e-10
            assert np.linalg.norm(g[k] - g0[k]) <= 1e-10 * np.linalg.norm(g0[k])


def test_batch_lanes_match(lib):
    """Groups overlap on several streams by default; one or two streams must give
    the same results (each group writes only its own buffers).  Groups on odd
    streams launch their classes in reverse order, so this also compares the
    two class orders."""
    systems = workloads.config3()
    pick = [2, 3, 51, 52, 93, 94]
    b = lib.Batch([systems[i] for i in pick])
    e0, g0 = b.run(want_grad=True)
    for n in (1, 2):
        b.set_lanes(n)
        for rep in range(2):                       # direct run, then a replay of the recaptured graph
            e, g = b.run(want_grad=True)
            for k in range(len(pick)):
                assert abs(e[k] - e0[k]) < 1e-10, (n, rep, k, e[k], e0[k])
                assert np.linalg.norm(g[k] - g0[k]) <= 1e-10 * np.linalg.norm(g0[k])


def test_pair_chunk_matches(lib):
    """Energy units of 64 bra pairs x a chunk of ket pairs (screened kets skipped
    inside the unit) visit exactly the quartets of one-ket units: the same
    quartet and skip counts, and the oracle's energy to 1e-9 Eh, for every width."""
    s = workloads.config1()
    g = CONFIGS["config1"]
    ref = None
    for kch in (1, 3, 16, 64):
        with lib.Context(s, pair_chunk=kch) as ctx:
            e = ctx.energy()
            st = ctx.counters()
        print(f"pair_chunk={kch}: gpu={e:.12f} quartets={st[0]:.0f} skipped={st[2]:.0f}")
        assert abs(e - g["energy"]) < 1e-9
        assert int(st[0]) + int(st[2]) == _unique_quartets(s)
        if ref is None:
            ref = (e, st[0], st[1], st[2])
        else:
            assert abs(e - ref[0]) < 1e-10
            assert (st[0], st[1], st[2]) == ref[1:]
    with pytest.raises(lib.VsvbError, match="pair_chunk"):
        lib.Context(s, pair_chunk=65)


def test_pair_chunk_gradient_matches(lib):
    """J/K units that claim several ket pairs (bra pairs when transposed) and
    skip the screened ones give the one-pair-claim gradient to 1e-10 relative,
    with the same quartet and skip counts."""
    s = workloads.config1()
    ref = None
    for kch in (1, 5, 64):
        with lib.Context(s, pair_chunk=kch) as ctx:
            e, g = ctx.energy_grad()
            st = ctx.counters()
        print(f"pair_chunk={kch}: E={e:.12f} |g|={np.linalg.norm(g):.10e} quartets={st[0]:.0f} skipped={st[2]:.0f}")
        assert int(st[0]) + int(st[2]) == _unique_quartets(s)
        if ref is None:
            ref = (e, g, st[0], st[1], st[2])
        else:
            assert abs(e - ref[0]) < 1e-9
            assert np.linalg.norm(g - ref[1]) <= 1e-10 * np.linalg.norm(ref[1])
            assert (st[0], st[1], st[2]) == ref[2:]


def test_batch_group_launch_matches(lib):
    """Class launches over many groups (persistent workgroups that drain one
    group's work units after another, one to three chunks per stream) give the
    results of one pass per group: a group of four rotations, displaced
    geometries, alkanes (two ethanes and two pentanes share a pass each) and a
    high-spin atom (two spin densities); direct runs and graph replays."""
    systems = workloads.config3()
    pick = [48, 49, 50, 51, 4, 5, 80, 84, 83, 90, 85]
    members = [systems[i] for i in pick] + [read_input(os.path.join(GOLDEN, "inputs", "testing__b"))]
    ref = lib.Batch(members)
    ref.set_group_launch(0)
    e0, g0 = ref.run(want_grad=True)
    st0 = ref.class_stats()
    ref.close()
    for chunks in (1, 2, 3):
        b = lib.Batch(members)
        b.set_group_launch(chunks)
        b.set_profiling(True)
        for rep in range(2):                       # direct run, then a graph replay
            e, g = b.run(want_grad=True)
            st = b.class_stats()
            for k in range(len(members)):
                assert abs(e[k] - e0[k]) < 1e-10, (chunks, rep, k, e[k], e0[k])
                assert np.linalg.norm(g[k] - g0[k]) <= 1e-10 * np.linalg.norm(g0[k])
            # the same quartets; one launch per chunk and kernel variant, timed
            assert np.array_equal(st[:, 1:4], st0[:, 1:4])
            ran = st[:, 1] > 0
            assert np.all(st[ran, 5] >= 1) and np.all(st[ran, 5] <= 4 * chunks * 2)
            assert np.all(st[ran, 0] > 0)
            print(f"chunks={chunks} rep={rep}: launches per class {st[:, 5].astype(int).tolist()}")
        assert b.uses_graph(), b.graph_error()
        b.close()


def test_batch_many_density_group(lib):
    """A group of more than four densities (six rotations of C8H18) runs the
    launch policies tuned for many densities; every member must still match its
    own independent evaluation.  The tuner picks only among launches below
    occupancy (they overlap better on the batch's four streams)."""
    systems = workloads.config3()
    pick = [52, 53, 54, 55, 56, 57, 3]
    b = lib.Batch([systems[i] for i in pick])
    e, g = b.run(want_grad=True)
    t = b.tuning()
    assert t[1, :, 7].sum() > 0          # the many-density bucket was tuned
    # launches stay below occupancy: 1-2 straight-line or 2-8 cooperative workgroups per CU
    for row in t.reshape(-1, 8)[t.reshape(-1, 8)[:, 7] > 0]:
        assert row[1] in ((2, 4, 8) if row[0] else (1, 2)), row
    e2, g2 = b.run(want_grad=True)       # replay of the captured graph
    for k, i in enumerate(pick):
        with lib.Context(systems[i]) as ctx:
            e1, g1 = ctx.energy_grad()
        print(f"wfn {i}: batch {e[k]:.12f} single {e1:.12f}")
        assert abs(e[k] - e1) < 1e-9 and abs(e2[k] - e1) < 1e-9
        assert np.linalg.norm(g[k] - g1) <= 1e-10 * np.linalg.norm(g1)
        assert np.linalg.norm(g2[k] - g1) <= 1e-10 * np.linalg.norm(g1)
