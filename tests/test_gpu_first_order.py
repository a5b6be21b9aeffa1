"""The reference's first-order orbital optimization (first_order_opt inside
minimize_energy, src/valence.F90:527-844, 2744-2885; valence_amd.optimize.
minimize_energy_first_order) on the library's energies and gradients, against
the converged totals the reference records for its first-order inputs
(testing/testing.py:69-150, tolerance 1e-6 Eh, :24): closed shells, high-spin
(unpaired orbitals; doubly occupied ones with spin averaging), NDF-expanded
orbitals, excited orbitals (xorb roots) and spin-coupled wave functions with
their spin-coupling optimization.  testing/c2-pt-vshf-p2 stops at max_iter =
10, so it also pins the iteration itself; testing/lih-sv optimizes a doubly
occupied orbital inside a spin-coupled wave function.
"""
import json
import os

import pytest

from conftest import GOLDEN
from valence_amd.system import read_input

pytestmark = pytest.mark.gpu

KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))

CASES = ["testing/be", "testing/he", "testing/h", "testing/h+ndf", "testing/be+ndf", "testing/h2o-vdz",
         "testing/be3s2", "testing/he3s-1s3s", "examples/li_opt", "testing/c2-pt-vshf-p2", "testing/cu+",
         "testing/h2-sz", "testing/h2-dz", "testing/be-sc", "testing/he1s2s", "testing/h2o-vdz-sc1",
         "testing/be-scv3s+2sc", "testing/lih-sv"]


def _case(name):
    from valence_amd import lib
    lib.load()
    grp, base = name.split("/")
    return read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{base}")), KNOWN[grp][base]


@pytest.mark.parametrize("name", CASES)
def test_first_order_matches_reference(name):
    from valence_amd.optimize import minimize_energy_first_order
    s, rec = _case(name)
    assert s.ptbnmax == 0 and len(s.orbset) > 0
    print(f"first-order {name}: orbset {s.orbset.tolist()} xorb {s.xorb.tolist()} nspinc {s.nspinc}", flush=True)
    r = minimize_energy_first_order(s)
    d = r.energy - rec["total_energy"]
    print(f"  E={r.energy:.12f} recorded={rec['total_energy']:.12f} diff={d:+.2e} iterations={r.iterations} "
          f"E(weights)={r.energy_at_weights:.12f}")
    assert abs(r.history[0] - rec["guess_energy"]) < 1e-8
    assert abs(d) < 1e-6
