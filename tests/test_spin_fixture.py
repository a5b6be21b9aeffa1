"""The oracle's spin-coupling matrices (tests/golden/spin_matrices.json) against
the reference's recorded answers: at each input's recorded spin-coupling
weights, c.H.c / c.S.c + E_nuc is the recorded guess energy (testing/testing.py,
examples/test_examples.py; tolerance 1e-10 Eh), and S is symmetric positive definite."""
import json
import os

import numpy as np

from conftest import GOLDEN

MATS = json.load(open(os.path.join(GOLDEN, "spin_matrices.json")))
KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))


def test_fixture_reproduces_recorded_energies():
    assert len(MATS) == 5
    for name, m in MATS.items():
        grp, case = name.split("__")
        H, S, c = np.array(m["ham"]), np.array(m["ovl"]), np.array(m["coeff_sc"])
        e = c @ H @ c / (c @ S @ c) + m["e_nuc"]
        assert abs(e - KNOWN[grp][case]["guess_energy"]) < 1e-10, name
        assert np.array_equal(H, H.T) and np.array_equal(S, S.T)
        assert np.linalg.eigvalsh(S).min() > 0
