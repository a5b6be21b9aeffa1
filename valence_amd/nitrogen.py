"""The NITROGEN potential-surface plug-in over the HIP library, and batched
finite-difference Hessians / harmonic frequencies.

The reference exposes VSVB to NITROGEN through four Fortran entry points
(src/valence_api_nitrogen.F90): init (:3-6), getn (:10-17), calcsurface
(:22-30: energy in cm^-1 at Cartesian coordinates in angstrom) and finalize
(:33-35).  NITROGEN then calls calcsurface once per geometry of its
finite-difference stencils (nitrogen/h2o/VSVB_STAT.job: GRAD_STEP 1e-3,
FD_ORDER 2, CALC_FREQ).  `Surface` mirrors the four entry points over
`vsvb_surface`; `stencil_hessian` evaluates a whole central-difference stencil in
one batch run (`vsvb_surface_batch`), every displaced geometry its own group
on the batch's overlapping streams.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import lib
from .system import ANGSTROM_TO_BOHR, VsvbSystem

HARTREE_TO_WAVENUMBER = 219474.631        # src/valence_api_nitrogen.F90:28
AMU_TO_ELECTRON_MASS = 1822.888486209
# most abundant isotope masses (u) by nuclear charge, as NITROGEN's MASS keyword uses
ISOTOPE_MASS = {1: 1.00782503207, 2: 4.00260325415, 3: 7.016004548, 4: 9.012182201, 5: 11.009305406,
                6: 12.0, 7: 14.00307400478, 8: 15.99491461956, 9: 18.998403224, 10: 19.99244017542}


class Surface:
    """init / getn / calcsurface / finalize of the reference's NITROGEN plug-in."""

    def __init__(self, system: VsvbSystem, device: int = 0):        # init
        self.system = system
        self._ctx = lib.Context(system, device=device)

    def getn(self) -> int:
        return int(self.system.natom)

    def calcsurface(self, x_angstrom) -> float:
        """Energy (cm^-1) at Cartesian coordinates [natom, 3] in angstrom."""
        return self._ctx.surface(np.asarray(x_angstrom, float))

    def finalize(self):
        self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.finalize()


def stencil(x0_angstrom, h: float = 1e-3):
    """Central-difference stencil for a Hessian in Cartesian coordinates (angstrom):
    x0, x0 +- h e_i (i < 3N), and x0 +- h e_i +- h e_j (i < j).  Returns the
    geometries and, for each, its (i, si, j, sj) displacement label."""
    x0 = np.asarray(x0_angstrom, float).reshape(-1)
    n = x0.size
    geoms, labels = [x0.copy()], [(None, 0, None, 0)]
    for i in range(n):
        for si in (1, -1):
            g = x0.copy(); g[i] += si * h
            geoms.append(g); labels.append((i, si, None, 0))
    for i in range(n):
        for j in range(i + 1, n):
            for si in (1, -1):
                for sj in (1, -1):
                    g = x0.copy(); g[i] += si * h; g[j] += sj * h
                    geoms.append(g); labels.append((i, si, j, sj))
    return [g.reshape(-1, 3) for g in geoms], labels


def stencil_hessian(system: VsvbSystem, x0_angstrom, h: float = 1e-3, device: int = 0,
                    batch: Optional[lib.Batch] = None):
    """Energy (Eh) and Cartesian Hessian (Eh / bohr^2) of the surface at x0
    (angstrom) by central differences with step h (angstrom), every stencil
    geometry in one batch run.  Pass `batch` (made by `stencil_batch`) to reuse
    its contexts for another point with the same stencil size."""
    geoms, labels = stencil(x0_angstrom, h)
    if batch is None:
        batch = stencil_batch(system, geoms, device)
    v = batch.surface(geoms) / HARTREE_TO_WAVENUMBER
    n = geoms[0].size
    e0 = v[0]
    ep = {}
    for k, (i, si, j, sj) in enumerate(labels):
        ep[(i, si, j, sj)] = v[k]
    H = np.zeros((n, n))
    hb = h * ANGSTROM_TO_BOHR
    for i in range(n):
        H[i, i] = (ep[(i, 1, None, 0)] - 2.0 * e0 + ep[(i, -1, None, 0)]) / hb ** 2
        for j in range(i + 1, n):
            H[i, j] = H[j, i] = (ep[(i, 1, j, 1)] - ep[(i, 1, j, -1)] - ep[(i, -1, j, 1)] + ep[(i, -1, j, -1)]) / (
                4.0 * hb ** 2)
    return e0, H


def stencil_batch(system: VsvbSystem, geoms: Sequence[np.ndarray], device: int = 0) -> lib.Batch:
    """A batch with one wave function per stencil geometry (each its own group)."""
    systems = [dataclasses.replace(system, coords=np.asarray(g, float) * ANGSTROM_TO_BOHR) for g in geoms]
    return lib.Batch(systems, device=device)


def harmonic_frequencies(system: VsvbSystem, hessian: np.ndarray, masses: Optional[List[float]] = None):
    """Harmonic frequencies (cm^-1, imaginary ones negative) of a Cartesian
    Hessian in Eh / bohr^2: eigenvalues of the mass-weighted Hessian."""
    if masses is None:
        masses = [ISOTOPE_MASS[int(round(system.nuc_charge[t]))] for t in system.atom_type]
    m = np.repeat(np.asarray(masses, float) * AMU_TO_ELECTRON_MASS, 3)
    w = np.linalg.eigvalsh(hessian / np.sqrt(np.outer(m, m)))
    return np.sign(w) * np.sqrt(np.abs(w)) * HARTREE_TO_WAVENUMBER
