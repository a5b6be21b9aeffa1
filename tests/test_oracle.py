"""Pins of the CPU oracle (the C restatement the HIP path is checked against).

Each check is independent of the oracle's own formulas: closed forms written
out here, mpmath at 50 digits, library routines (numpy), the reference's own
compiled Givens kernel (oracle/_ref, built from src/givens.F90 and
src/givens_in_c.cpp when /root/reference is present), and the reference's
recorded energies (testing/testing.py:69-150, examples/test_examples.py:63-152).
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from valence_amd.system import read_input

KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))
ORACLE_KNOWN = json.load(open(os.path.join(GOLDEN, "oracle_known_answers.json")))
mp.mp.dps = 50


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def boys_mp(n, T):
    T = mp.mpf(T)
    return mp.quad(lambda t: t ** (2 * n) * mp.e ** (-T * t * t), [0, 1])


@pytest.mark.parametrize("T", [0.0, 1e-9, 0.3, 2.5, 11.0, 29.99, 30.01, 45.0, 250.0])
def test_boys_against_quadrature(T):
    F = oracle.boys(10, T)
    for n in range(11):
        ref = float(boys_mp(n, T))
        assert abs(F[n] - ref) <= 2e-15 * ref, (n, T, F[n], ref)


def test_norm_prim_closed_form():
    """Normalized contraction: sum c_i c_j <x^l g_i|x^l g_j> = 1 with the
    one-dimensional Gaussian moment int x^{2l} e^{-a x^2} = (2l-1)!! / (2a)^l sqrt(pi/a)."""
    rng = np.random.default_rng(5)
    for l in (0, 1, 2):
        e = np.sort(np.exp(rng.uniform(-2, 8, 5)))[::-1]
        u = rng.uniform(0.1, 1, 5)
        c = oracle.norm_prim(l, e, u)
        df = float(np.prod(np.arange(1, 2 * l, 2))) if l > 0 else 1.0
        s = 0.0
        for i in range(5):
            for j in range(5):
                a = e[i] + e[j]
                s += c[i] * c[j] * (math.pi / a) ** 1.5 * df / (2 * a) ** l
        assert abs(s - 1.0) < 1e-13


def _ssss(a, A, b, B, c, C, d, D):
    p, q = a + b, c + d
    P = (a * np.array(A) + b * np.array(B)) / p
    Q = (c * np.array(C) + d * np.array(D)) / q
    rho = p * q / (p + q)
    K = mp.e ** (-a * b / p * float(np.sum((np.array(A) - B) ** 2)) - c * d / q * float(np.sum((np.array(C) - D) ** 2)))
    T = rho * float(np.sum((P - Q) ** 2))
    F0 = mp.mpf(1) if T == 0 else mp.sqrt(mp.pi / (4 * T)) * mp.erf(mp.sqrt(T))
    return float(2 * mp.pi ** 2.5 / (p * q * mp.sqrt(p + q)) * K * F0)


def test_eri_ssss_closed_form():
    rng = np.random.default_rng(7)
    for _ in range(20):
        ex = np.exp(rng.uniform(-2, 6, 4))
        cen = rng.normal(0, 1.5, (4, 3))
        blk = oracle.eri_block([(cen[k], 0, [ex[k]], [1.0]) for k in range(4)])
        ref = _ssss(ex[0], cen[0], ex[1], cen[1], ex[2], cen[2], ex[3], cen[3])
        # exp() of an argument x carries a relative error ~ |x| eps; arguments here reach ~200
        assert abs(blk[0] - ref) <= 1e-13 * abs(ref) + 1e-300


def _rand_shell(rng, l, nprim):
    return (rng.normal(0, 1.0, 3), l, np.exp(rng.uniform(-1, 3, nprim)), rng.uniform(0.2, 1, nprim))


def _block(shells):
    dims = [(s[1] + 1) * (s[1] + 2) // 2 for s in shells]
    return oracle.eri_block(shells).reshape(dims)


@pytest.mark.parametrize("ls", [(1, 0, 2, 1), (2, 2, 1, 0), (2, 1, 2, 2), (1, 1, 1, 1)])
def test_eri_permutational_symmetry(ls):
    rng = np.random.default_rng(sum(ls))
    sh = [_rand_shell(rng, l, 2) for l in ls]
    a, b, c, d = sh
    g = _block([a, b, c, d])
    assert np.allclose(_block([b, a, c, d]), g.transpose(1, 0, 2, 3), rtol=1e-12, atol=1e-14)
    assert np.allclose(_block([a, b, d, c]), g.transpose(0, 1, 3, 2), rtol=1e-12, atol=1e-14)
    assert np.allclose(_block([c, d, a, b]), g.transpose(2, 3, 0, 1), rtol=1e-12, atol=1e-14)


def test_eri_translation_invariance():
    rng = np.random.default_rng(3)
    sh = [_rand_shell(rng, l, 2) for l in (2, 1, 1, 0)]
    g = _block(sh)
    shift = np.array([0.7, -1.3, 2.1])
    g2 = _block([(s[0] + shift, s[1], s[2], s[3]) for s in sh])
    assert np.allclose(g, g2, rtol=1e-12, atol=1e-14)


def test_eri_derivative_identity():
    """d/dA_x (s s|s s) = 2a (p_x s|s s) for a single primitive (Gaussian derivative rule)."""
    rng = np.random.default_rng(11)
    a = 1.3
    A = rng.normal(0, 1, 3)
    others = [_rand_shell(rng, 0, 1) for _ in range(3)]
    h = 1e-4
    Ap, Am = A.copy(), A.copy()
    Ap[0] += h
    Am[0] -= h
    fp = _block([(Ap, 0, [a], [1.0])] + others).ravel()[0]
    fm = _block([(Am, 0, [a], [1.0])] + others).ravel()[0]
    px = _block([(A, 1, [a], [1.0])] + others).reshape(3, 1, 1, 1)[0, 0, 0, 0]
    assert abs((fp - fm) / (2 * h) - 2 * a * px) < 1e-7 * max(1.0, abs(px))


def test_givens_det_vs_numpy():
    rng = np.random.default_rng(2)
    for n in (2, 5, 9, 17):
        a = rng.normal(size=(n, n)) + n * np.eye(n)
        assert abs(oracle.givens_det(a, 0.0) - np.linalg.det(a)) <= 1e-12 * abs(np.linalg.det(a))


def test_givens_skip_semantics():
    """A rotation is skipped when a(j,i-1)^2 + a(j,i)^2 <= tol (givens.F90:245): with a
    huge tol nothing rotates and the 'determinant' is the plain diagonal product."""
    rng = np.random.default_rng(4)
    a = rng.normal(size=(6, 6))
    assert oracle.givens_det(a, 1e300) == pytest.approx(np.prod(np.diag(a)), rel=0, abs=0)


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference sources not present")
def test_givens_matches_reference_build():
    import subprocess
    subprocess.run(["make", "-C", oracle.HERE, "ref"], check=True, capture_output=True)
    rng = np.random.default_rng(9)
    for n in (3, 8, 33):
        for tol in (0.0, 1e-16, 1e-3):
            a = rng.normal(size=(n, n))
            mine = oracle.givens_det(a, tol)
            for variant in ("fortran", "c"):
                ref = oracle.ref_givens_det(a, tol, variant)
                assert abs(mine - ref) <= 1e-13 * abs(ref), (n, tol, variant)


FAST = sorted(k for k, v in ORACLE_KNOWN.items() if v["seconds"] < 0.5)


@pytest.mark.parametrize("key", FAST)
def test_known_answers(key):
    """Guess energies recorded by the reference, at its own tolerance 1e-8 Eh (testing.py:23)."""
    grp, name = key.split("/")
    s = read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{name}"))
    e = oracle.energy(s, nthreads=4).energy
    assert abs(e - KNOWN[grp][name]["guess_energy"]) < 1e-8
    # and the oracle is stable against its own recorded value
    assert abs(e - ORACLE_KNOWN[key]["oracle_energy"]) < 1e-10


def test_known_answer_record_is_complete():
    """The committed record covers all 58 reference answers and all pass at 1e-8."""
    assert len(ORACLE_KNOWN) == 58
    assert max(abs(v["diff"]) for v in ORACLE_KNOWN.values()) < 1e-8


@pytest.mark.parametrize("name", ["testing__h2o-vdz", "testing__be-sc", "testing__lih"])
def test_direct_int2e_equals_table(name):
    """The orbital ERI table (AO half-transform) reproduces the literal int2e loop."""
    s = read_input(os.path.join(GOLDEN, "inputs", name))
    a = oracle.energy(s, eri_table=True).energy
    b = oracle.energy(s, eri_table=False).energy
    assert abs(a - b) < 1e-11


@pytest.mark.parametrize("plant", [1, 2])
def test_planted_mistake_is_caught(plant):
    """Plant a plausible mistake in the oracle (exchange term dropped / sign flipped):
    the known-answer check must reject it by far more than its 1e-8 tolerance."""
    for name in ("h2o-vdz", "be", "lih"):
        s = read_input(os.path.join(GOLDEN, "inputs", f"testing__{name}"))
        e = oracle.energy(s, plant=plant).energy
        assert abs(e - KNOWN["testing"][name]["guess_energy"]) > 1e-3
