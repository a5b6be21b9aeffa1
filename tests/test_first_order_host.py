"""Host side of the first-order orbital step (valence_amd/optimize.py), on the CPU.

* The pencil recovery: for a Rayleigh quotient E(x) = x^T H x / x^T S x of a
  known symmetric H and positive-definite S, the energies and exact gradients
  at n + 4 points give (H, S) up to one factor -- the same eigenvalues and
  eigenvectors as scipy's generalized eigensolver on the true pencil, and a
  planted sign error in the gradient formula is detected.
* The split system of a doubly occupied orbital: the replacement holds the
  alpha electron as the last unpaired orbital and the original the beta one as
  the last doubly occupied orbital (the layout vsvb_options.beta_only_last
  describes); orbitals, atoms, weights and NDF references stay consistent.
"""
import os

import numpy as np
import pytest
from scipy.linalg import eigh

from conftest import GOLDEN
from valence_amd.optimize import _pencil_from_gradients, _substituted
from valence_amd.system import read_input


def _pencil(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    S = A @ A.T + n * np.eye(n)
    B = rng.standard_normal((n, n))
    H = -(B + B.T) - 3.0 * S
    return H, S


@pytest.mark.parametrize("n", [1, 2, 5, 13])
def test_pencil_from_exact_gradients(n):
    H, S = _pencil(n, 10 + n)

    def ev(x):
        d = x @ S @ x
        e = (x @ H @ x) / d
        return e, 2.0 * (H @ x - e * S @ x) / d

    c = np.random.default_rng(1).standard_normal(n)
    Hr, Sr = _pencil_from_gradients(ev, c)
    f = (c @ S @ c)
    assert np.allclose(Hr * f, H, rtol=1e-9, atol=1e-9 * np.abs(H).max())
    assert np.allclose(Sr * f, S, rtol=1e-9, atol=1e-9 * np.abs(S).max())
    w0, v0 = eigh(H, S)
    w1, v1 = eigh(Hr, Sr)
    assert np.allclose(w0, w1, rtol=1e-10, atol=1e-10)
    for k in range(n):
        assert abs(abs(v0[:, k] @ S @ v1[:, k]) / np.sqrt(v1[:, k] @ S @ v1[:, k]) - 1.0) < 1e-8


def test_pencil_rejects_inconsistent_gradients():
    H, S = _pencil(6, 3)

    def ev(x):  # planted mistake: the sign of the S x term
        d = x @ S @ x
        e = (x @ H @ x) / d
        return e, 2.0 * (H @ x + e * S @ x) / d

    with pytest.raises(RuntimeError, match="not determined"):
        _pencil_from_gradients(ev, np.ones(6))


@pytest.mark.parametrize("beta", [False, True])
@pytest.mark.parametrize("deck,k", [("testing__c2-pt-vshf-p2", 2), ("testing__lih-sv", 0)])
def test_split_system_layout(beta, deck, k):
    # c2-pt-vshf-p2: unpaired, doubly occupied and NDF orbitals; lih-sv: a spin-coupled pair first
    s = read_input(os.path.join(GOLDEN, "inputs", deck))
    nnd = 2 * s.npair + s.nunpd
    p = nnd + k                                             # doubly occupied orbital k
    t, assemble, xs = _substituted(s, p, beta)
    assert t.nunpd == s.nunpd + 1 and t.ndocc == s.ndocc and t.norbs == s.norbs + 1 and t.npair == s.npair
    t.coeff = assemble(s.coeff, np.arange(xs.stop - xs.start, dtype=float) + 100.0)
    t.validate()
    for q in range(2 * s.npair):                            # spin-coupled orbitals keep their place
        a = slice(int(s.orb_start[q]), int(s.orb_start[q + 1]))
        b = slice(int(t.orb_start[q]), int(t.orb_start[q + 1]))
        assert np.array_equal(s.coeff[a], t.coeff[b])
    sl = slice(int(s.orb_start[p]), int(s.orb_start[p + 1]))
    x_pos = nnd + s.ndocc if beta else nnd                  # the replacement's orbital index in t
    p_pos = nnd if beta else nnd + s.ndocc                  # the original's
    for q, w in ((x_pos, 100.0 + np.arange(sl.stop - sl.start)), (p_pos, s.coeff[sl])):
        tq = slice(int(t.orb_start[q]), int(t.orb_start[q + 1]))
        assert np.array_equal(t.coeff[tq], w)
        assert np.array_equal(t.xpset[tq], s.xpset[sl])
        assert t.orb_natom[q] == s.orb_natom[p] and np.array_equal(t.orb_atoms[q], s.orb_atoms[p])
    assert xs == slice(int(t.orb_start[x_pos]), int(t.orb_start[x_pos + 1]))
    # the NDFs keep their weights and their place at the end; references to them are unchanged
    for f in range(s.ndf):
        a = slice(int(s.orb_start[s.norbs - s.ndf + f]), int(s.orb_start[s.norbs - s.ndf + f + 1]))
        b = slice(int(t.orb_start[t.norbs - t.ndf + f]), int(t.orb_start[t.norbs - t.ndf + f + 1]))
        assert np.array_equal(s.coeff[a], t.coeff[b]) and np.array_equal(s.xpset[a], t.xpset[b])
