"""The 'orbitals' / 'nelecwfn' writer (valence_amd.system.format_orbitals,
format_nelecwfn) against the reference's own files: two of its test decks end
with the text xm_output wrote (src/xm_module.F90:464-577) --
testing/test_cases/he1s2s (a finished run: mode 'done') and
testing/test_cases/be-scv3s+2sc (two spin couplings and an excited-orbital
root: the nelecwfn records, then the orbitals of a 'save').  The writer must
reproduce those records character for character from the decks' own
orbitals, and its output must read back as the orbital part of an input deck.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd.system import (format_nelecwfn, format_orbitals, parse_input, read_input, write_input)


def _deck(name):
    path = os.path.join(GOLDEN, "inputs", name)
    return read_input(path), open(path).read().split("\n")


def _block(lines, first, last):
    i = lines.index(first)
    j = next(k for k in range(i, len(lines)) if lines[k].startswith(last))
    return lines[i:j + 1]


def test_orbitals_he1s2s_done():
    s, lines = _deck("testing__he1s2s")
    want = _block(lines, "      1        1   7", " converged to")
    got = format_orbitals(s, -2.1365129715011482, mode="done", etol=1e-6).split("\n")
    assert got[:len(want)] == want
    assert got[len(want):] == ["", ""]  # the trailing blank record


def test_nelecwfn_and_orbitals_be_scv3s_2sc():
    s, lines = _deck("testing__be-scv3s+2sc")
    assert s.nspinc == 2 and len(s.xorb) == 1
    want_n = _block(lines, "    0.37391091  1  2    3  4", "   2  1") + [""]
    assert format_nelecwfn(s).split("\n")[:-1] == want_n
    want_o = _block(lines, "      1        1  11", " total energy in atomic units")
    got = format_orbitals(s, -14.3173525131288084).split("\n")
    assert got[:len(want_o)] == want_o and got[len(want_o):] == ["", ""]


def test_single_coupling_has_no_nelecwfn():
    s, _ = _deck("testing__h2o-vdz-sc1")
    assert s.nspinc <= 1 and format_nelecwfn(s) == ""


@pytest.mark.parametrize("name", ["testing__ethane", "testing__be+ndf", "testing__be-scv3s+2sc", "testing__cu+"])
def test_orbitals_read_back(name):
    """Input deck without orbitals + the orbitals file = the deck, with the
    weights at the file's eight decimals (long headers wrap onto new records)."""
    s, _ = _deck(name)
    rng = np.random.default_rng(7)
    s.coeff = s.coeff + 1e-3 * rng.standard_normal(s.coeff.shape)
    text = write_input(s, orbitals=False) + format_orbitals(s, -1.0)
    r = parse_input(text)
    assert np.array_equal(r.xpset, s.xpset) and np.array_equal(r.orb_start, s.orb_start)
    assert np.array_equal(r.orb_natom, s.orb_natom) and np.array_equal(r.orb_atoms, s.orb_atoms)
    assert np.max(np.abs(r.coeff - np.round(s.coeff, 8))) < 1e-12


def test_header_wraps_like_fortran():
    """Format (5x,1i2,5x,10i4) with 12 items: the eleventh item closes the first
    record; reversion starts a new record with the i2 field."""
    from valence_amd.system import _fmt_header, _fe
    lines = _fmt_header([11] + list(range(1, 12)) + [40])
    assert lines == ["     11" + "     " + "".join(f"{k:4d}" for k in range(1, 11)), "     11       40"]
    assert _fe(1e-6, 10, 2) == "  0.10E-05" and _fe(0.0996, 10, 2) == "  0.10E+00" and _fe(0.0, 10, 2) == "  0.00E+00"
