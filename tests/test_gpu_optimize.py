"""Orbital optimization on the GPU energies (valence_amd/optimize.py) against the
reference's recorded converged energies (testing/testing.py:69-150), at the
reference's own tolerance for converged totals, 1e-6 Eh (testing/testing.py:24).
The inputs are the reference's test inputs (tests/golden/inputs/testing__*),
with their optimization groups and method as given.

* h2o-vdz-dem (:94) and li_opt (:151) ask for the direct energy minimization
  (ptbnmax > 0): dem_minimize, the reference's DEM restated on the library's
  energies, must reproduce the recorded totals.
* h2o-vdz (:91) asks for the first-order method, which converges to the
  variational minimum over the group weights: the gradient-driven L-BFGS
  (minimize_energy, the consumer of the adjoint gradient) must reach the
  recorded total.  On the DEM inputs it must end at or below the DEM totals:
  equal for h2o-vdz-dem, 5.0e-5 Eh lower for li_opt, whose single-step DEM
  (perturbation 0.1, feather 0) stops above the minimum.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd.system import read_input

pytestmark = pytest.mark.gpu

KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))


def _case(name):
    from valence_amd import lib
    lib.load()
    s = read_input(os.path.join(GOLDEN, "inputs", f"testing__{name}"))
    return s, KNOWN["testing"][name]


@pytest.mark.parametrize("name", ["h2o-vdz-dem", "li_opt"])
def test_dem_matches_reference(name, tmp_path):
    from valence_amd.optimize import dem_minimize
    s, rec = _case(name)
    assert s.ptbnmax > 0
    print(f"DEM {name}: groups {s.orbset.tolist()} ptbnmax {s.ptbnmax} feather {s.feather}", flush=True)
    r = dem_minimize(s)
    print(f"  E={r.energy:.12f} recorded={rec['total_energy']:.12f} diff={r.energy - rec['total_energy']:+.2e} "
          f"energy calls={r.iterations}")
    assert abs(r.history[0] - rec["guess_energy"]) < 1e-8
    assert abs(r.energy - rec["total_energy"]) < 1e-6
    # the 'orbitals' file continues the run: read back into the deck, its orbitals
    # give the energy of the returned weights (to the file's eight decimals); the
    # reference's restore after two failed trials leaves that above the tracked
    # total (h2o-vdz-dem: 5.8e-4 Eh), the file reports the tracked total
    from valence_amd import lib
    from valence_amd.optimize import save_orbitals
    from valence_amd.system import parse_input, write_input
    save_orbitals(s, r, str(tmp_path))
    text = open(tmp_path / "orbitals").read()
    assert " converged to " in text
    s2 = parse_input(write_input(s, orbitals=False) + text)
    with lib.Context(s) as ctx:
        e_direct = ctx.energy(r.coeff)
        e_round = ctx.energy(np.round(r.coeff, 8))
    with lib.Context(s2) as ctx:
        e2 = ctx.energy()
    print(f"  at the final weights: E={e_direct:.12f}; rounded to 8 decimals: E={e_round:.12f}; "
          f"from the orbitals file: E={e2:.12f}; |c| in [{np.min(np.abs(r.coeff[r.coeff != 0])):.3e}, "
          f"{np.max(np.abs(r.coeff)):.3e}]; max |file - c| {np.max(np.abs(s2.coeff - r.coeff)):.2e}")
    assert abs(e_direct - r.energy_at_weights) < 1e-9
    assert abs(e2 - r.energy_at_weights) < 1e-6
    assert f"{r.energy:32.16f}" in text


@pytest.mark.parametrize("name", ["h2o-vdz", "h2o-vdz-dem", "li_opt"])
def test_gradient_optimizer(name):
    from valence_amd.optimize import minimize_energy
    s, rec = _case(name)
    print(f"L-BFGS {name}: groups {s.orbset.tolist()} guess {rec['guess_energy']:.12f}", flush=True)
    r = minimize_energy(s)
    d = r.energy - rec["total_energy"]
    print(f"  E={r.energy:.12f} recorded={rec['total_energy']:.12f} diff={d:+.2e} "
          f"iterations={r.iterations} |g|={r.gradient_norm:.2e}")
    assert abs(r.history[0] - rec["guess_energy"]) < 1e-8
    assert r.energy <= r.history[0] + 1e-12
    if name == "li_opt":
        assert -6e-5 < d < -4e-5      # below the single-step DEM's stopping point
    else:
        assert abs(d) < 1e-6
