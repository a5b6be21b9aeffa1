"""Geometries of the reference's NITROGEN example (nitrogen/h2o): the first
potential-surface call of its STAT job (VSVB_STAT.job: Z-matrix H1 / O 1 R12 /
H2 2 R23 1 A123, PESPERM 2 1 3, REF1 = 1.0 1.0 D110, GRAD_STEP 1e-3) is at
R12 - GRAD_STEP; its output (nitrogen/h2o/output:137) prints the guess energy
of the input wave function there."""
import math
import os

import numpy as np

from conftest import GOLDEN

H2O_INPUT = os.path.join(GOLDEN, "inputs", "nitrogen__h2o.inp")
H2O_FIRST_GUESS_ENERGY = -71.1251116551235896   # nitrogen/h2o/output:137


def h2o_geometry(r12=1.0, r23=1.0, a123_deg=110.0):
    """VALENCE atom order O, H1, H2 (PESPERM 2 1 3), angstrom; O at the origin,
    H1 on x, H2 in the xy plane."""
    a = math.radians(a123_deg)
    return np.array([[0.0, 0.0, 0.0], [r12, 0.0, 0.0], [r23 * math.cos(a), r23 * math.sin(a), 0.0]])


H2O_FIRST_POINT = h2o_geometry(1.0 - 1e-3)
