"""Spin-coupling matrices and the spin-optimization step on the GPU (through the C ABI).

References: tests/golden/spin_matrices.json, built by the oracle from one run per
unit weight vector and per pair of them (tests/golden/make_spin_matrices.py);
at the inputs' recorded weights those matrices reproduce the reference's
recorded guess energies to 4e-12 Eh.  Tolerances: matrix elements 1e-11 of the
largest element; energies 1e-9 Eh; the optimized weights against
scipy.linalg.eigh on the fixture matrices (|cos| >= 1 - 1e-10, c.S.c = 1 to 1e-10).
"""
import dataclasses
import json
import os

import numpy as np
import pytest
import scipy.linalg

from conftest import GOLDEN
from valence_amd.system import read_input

pytestmark = pytest.mark.gpu

MATS = json.load(open(os.path.join(GOLDEN, "spin_matrices.json")))
NAMES = sorted(MATS)


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


def _system(name):
    return read_input(os.path.join(GOLDEN, "inputs", name))


@pytest.mark.parametrize("name", NAMES)
def test_spin_matrices(lib, name):
    print(f"case {name}", flush=True)
    ref = MATS[name]
    Hr, Sr = np.array(ref["ham"]), np.array(ref["ovl"])
    s = _system(name)
    with lib.Context(s) as ctx:
        H, S = ctx.spin_matrices()
        e = ctx.energy()
    dh = np.abs(H - Hr).max() / np.abs(Hr).max()
    ds = np.abs(S - Sr).max() / np.abs(Sr).max()
    c = np.asarray(s.coeff_sc, float)
    eq = c @ H @ c / (c @ S @ c) + ref["e_nuc"]
    print(f"  rel dH={dh:.2e} dS={ds:.2e} E={e:.12f} c.H.c/c.S.c+Enuc={eq:.12f}")
    assert np.allclose(H, H.T, rtol=0, atol=0) and np.allclose(S, S.T, rtol=0, atol=0)
    assert dh < 1e-11 and ds < 1e-11
    assert abs(e - eq) < 1e-9


@pytest.mark.parametrize("name", NAMES)
def test_energy_is_quadratic_in_weights(lib, name):
    """E(c) = c.H.c / c.S.c + E_nuc for arbitrary weights (fresh contexts)."""
    s = _system(name)
    with lib.Context(s) as ctx:
        H, S = ctx.spin_matrices()
    rng = np.random.default_rng(sum(map(ord, name)))
    for _ in range(2):
        c = rng.normal(size=s.nspinc)
        with lib.Context(dataclasses.replace(s, coeff_sc=c)) as ctx:
            e = ctx.energy()
        assert abs(e - (c @ H @ c / (c @ S @ c) + MATS[name]["e_nuc"])) < 1e-9


@pytest.mark.parametrize("name", NAMES)
def test_spin_opt(lib, name):
    ref = MATS[name]
    Hr, Sr = np.array(ref["ham"]), np.array(ref["ovl"])
    w, v = scipy.linalg.eigh(Hr, Sr)
    s = _system(name)
    with lib.Context(s) as ctx:
        e_guess = ctx.energy()
        e_opt, c = ctx.spin_opt()
        ctx.set_spin_coupling(c)
        e_set = ctx.energy()
    cos = abs(c @ Sr @ v[:, 0]) / np.sqrt((c @ Sr @ c) * (v[:, 0] @ Sr @ v[:, 0]))
    print(f"{name}: E_guess={e_guess:.12f} E_opt={e_opt:.12f} scipy={w[0] + ref['e_nuc']:.12f} "
          f"E(set)={e_set:.12f} c={c}")
    assert abs(e_opt - (w[0] + ref["e_nuc"])) < 1e-9
    assert cos > 1 - 1e-10
    assert abs(c @ Sr @ c - 1.0) < 1e-10
    assert abs(e_set - e_opt) < 1e-9
    assert e_opt <= e_guess + 1e-12
    assert c[np.argmax(np.abs(c))] > 0


def test_spin_entries_reject_closed_shell(lib):
    from valence_amd import workloads
    with lib.Context(workloads.config0()) as ctx:
        with pytest.raises(lib.VsvbError, match="no spin-coupled pairs"):
            ctx.spin_matrices()
