"""Host logic: the input parser (xm_getdims/xm_input semantics) and the workload generators."""
import glob
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd.system import InputError, nuclear_repulsion, parse_input, read_input, write_input
from valence_amd import workloads

INPUTS = sorted(glob.glob(os.path.join(GOLDEN, "inputs", "*")))
KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))


@pytest.mark.parametrize("path", INPUTS, ids=os.path.basename)
def test_parse_roundtrip(path):
    s = read_input(path)
    s2 = parse_input(write_input(s))
    assert np.array_equal(s.xpset, s2.xpset) and np.array_equal(s.coeff, s2.coeff)
    assert np.allclose(s.coords, s2.coords, rtol=0, atol=1e-14)
    assert np.array_equal(s.prim_exp, s2.prim_exp) and np.array_equal(s.pair_sc, s2.pair_sc)


@pytest.mark.parametrize("grp,name", [(g, n) for g in ("testing", "examples") for n in KNOWN[g]])
def test_nuclear_repulsion_known(grp, name):
    """E_nuc against the reference's recorded values (tolerance testing.py:22)."""
    s = read_input(os.path.join(GOLDEN, "inputs", f"{grp}__{name}"))
    assert abs(nuclear_repulsion(s) - KNOWN[grp][name]["nuclear_repulsion"]) < 1e-10 or \
        KNOWN[grp][name]["nuclear_repulsion"] == 0.0


def test_list_directed_semantics():
    # values may span records; the rest of the last record is skipped; '/' ends a READ
    txt = open(os.path.join(GOLDEN, "inputs", "testing__h")).read()
    lines = txt.splitlines()
    lines[0] = lines[0] + "   999 trailing junk"
    s = parse_input("\n".join(lines))
    assert s.natom == 1 and s.nunpd == 1


def test_bad_inputs_raise():
    txt = open(os.path.join(GOLDEN, "inputs", "testing__h")).read()
    with pytest.raises(InputError):
        parse_input(txt.replace("     1 1        4", "     1 1        9", 1))
    with pytest.raises(InputError):
        parse_input("1 1 0 1 0 4 4 0 4 6 0 0 1 0 1\n")


def test_workload_sizes():
    """Sizes of the seeded workloads (AOs, shells, orbitals, weights)."""
    expect = {0: (25, 5, 85), 1: (210, 33, 690), 2: (72, 12, 204), 4: (1536, 384, 5376)}
    for i, (nbf, nsp, totlen) in expect.items():
        s = workloads.CONFIGS[i]()
        assert (s.nbf(), s.nspatial, s.totlen) == (nbf, nsp, totlen)


def test_alkane_geometry():
    for n in (2, 5, 8):
        elems, xyz, bonds = workloads.alkane_geometry(n)
        assert elems.count("C") == n and elems.count("H") == 2 * n + 2
        d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
        np.fill_diagonal(d, 9.0)
        assert d.min() > 1.0899
        for (c, h) in bonds:
            assert abs(d[c, h] - 1.09) < 1e-12
        for k in range(n - 1):
            assert abs(d[k, k + 1] - 1.54) < 1e-12


def test_config3_composition():
    b = workloads.config3()
    assert len(b) == 96
    assert sum(s.title.startswith("C8H18 displaced") for s in b) == 48
    assert sum(s.title.startswith("C8H18 rotated") for s in b) == 32
