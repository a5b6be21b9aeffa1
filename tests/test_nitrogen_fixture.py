"""The NITROGEN H2O fixture against the reference's own output: the oracle's
energy of the example's input wave function at the first potential-surface
point of its job equals the guess energy the reference printed there
(nitrogen/h2o/output:137), to 1e-10 Eh."""
import dataclasses

from nitrogen_cases import H2O_FIRST_GUESS_ENERGY, H2O_FIRST_POINT, H2O_INPUT
from valence_amd.system import ANGSTROM_TO_BOHR, read_input


def test_h2o_first_surface_point():
    import oracle
    oracle.build()
    s = read_input(H2O_INPUT)
    e = oracle.energy(dataclasses.replace(s, coords=H2O_FIRST_POINT * ANGSTROM_TO_BOHR)).energy
    assert abs(e - H2O_FIRST_GUESS_ENERGY) < 1e-10
