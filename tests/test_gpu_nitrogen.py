"""The NITROGEN plug-in path on the GPU (vsvb_surface, vsvb_surface_batch,
vsvb_batch_set_coords) and the batched stencil Hessian.

Reference value: the guess energy the reference printed at the first surface
point of its H2O example (nitrogen/h2o/output:137), 1e-8 Eh (the recorded
answers' tolerance).  Batched energies equal single evaluations to 1e-9 Eh; the
finite-difference Hessian is symmetric and obeys translation invariance
(sum over atoms of each Cartesian block = 0) to 1e-5 of its largest element
(finite-difference noise at h = 1e-3 angstrom).
"""
import dataclasses

import numpy as np
import pytest

from nitrogen_cases import H2O_FIRST_GUESS_ENERGY, H2O_FIRST_POINT, H2O_INPUT, h2o_geometry
from valence_amd import nitrogen
from valence_amd.system import ANGSTROM_TO_BOHR, read_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


def test_calcsurface_first_point(lib):
    s = read_input(H2O_INPUT)
    with nitrogen.Surface(s) as surf:
        assert surf.getn() == 3
        v = surf.calcsurface(H2O_FIRST_POINT)
    e = v / nitrogen.HARTREE_TO_WAVENUMBER
    print(f"v={v:.6f} cm-1  E={e:.13f}  reference {H2O_FIRST_GUESS_ENERGY:.13f}  diff {e - H2O_FIRST_GUESS_ENERGY:.2e}")
    assert abs(e - H2O_FIRST_GUESS_ENERGY) < 1e-8


def test_surface_batch_matches_single(lib):
    s = read_input(H2O_INPUT)
    geoms = [h2o_geometry(1.0 + 0.01 * k, 1.0 - 0.007 * k, 110.0 + k) for k in range(7)]
    b = nitrogen.stencil_batch(s, geoms)
    shifted = [g + np.array([0.2, -0.1, 0.05]) for g in geoms]      # new coordinates: a rigid shift
    for x in (geoms, shifted, geoms):
        v = b.surface(x)
        with lib.Context(s) as ctx:
            ref = np.array([ctx.surface(g) for g in x])
        assert np.abs(v - ref).max() / nitrogen.HARTREE_TO_WAVENUMBER < 1e-9
    assert np.abs(b.surface(shifted) - b.surface(geoms)).max() / nitrogen.HARTREE_TO_WAVENUMBER < 1e-9


def test_set_coords_keeps_groups(lib):
    """Wave functions created on one geometry share a quartet pass: they cannot be split."""
    s = read_input(H2O_INPUT)
    b = lib.Batch([s, s])
    x = s.coords.copy()
    with pytest.raises(lib.VsvbError, match="same coordinates"):
        b.set_coords([x, x + 0.1])
    b.set_coords([x + 0.1, x + 0.1])
    e, _ = b.run(want_grad=False)
    with lib.Context(dataclasses.replace(s, coords=x + 0.1)) as ctx:
        assert abs(e[0] - ctx.energy()) < 1e-9 and abs(e[1] - e[0]) < 1e-12


def test_stencil_hessian(lib):
    s = read_input(H2O_INPUT)
    e0, H = nitrogen.stencil_hessian(s, H2O_FIRST_POINT, h=1e-3)
    assert abs(e0 - H2O_FIRST_GUESS_ENERGY) < 1e-8
    big = np.abs(H).max()
    assert np.abs(H - H.T).max() <= 1e-12 * big
    for a in range(3):                      # translation invariance per Cartesian direction pair
        for b_ in range(3):
            block = H[a::3, b_::3]
            assert np.abs(block.sum(axis=1)).max() < 1e-5 * big, (a, b_)
    freq = nitrogen.harmonic_frequencies(s, H)
    print("harmonic frequencies (cm-1):", np.round(freq, 1))
    assert np.isfinite(freq).all()
