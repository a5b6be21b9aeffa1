"""The allocation-free entry point, vsvb_energy_grad_static, through the C ABI.

Two calls in a row, the second on a smaller system, must give what
vsvb_energy_grad gives on a context made from the same system (energies to
1e-10 Eh, gradients to 1e-10 norm-wise relative: the same kernels on the same
inputs, differing only in the order of floating-point atomics), and the
oracle's energies to 1e-9 Eh.  A system larger than the static scratch array
and a spin-coupled one are refused with clear errors.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd import workloads

pytestmark = pytest.mark.gpu

CONFIGS = json.load(open(os.path.join(GOLDEN, "config_energies.json")))


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


def test_static_entry_matches_context_entry(lib):
    cases = [("config1", workloads.config1()), ("config0", workloads.config0())]
    static = []
    for name, s in cases:                      # two calls in a row, the second smaller
        print(f"static entry: {name} ({s.title})", flush=True)
        static.append(lib.energy_grad_static(s))
    again = lib.energy_grad_static(cases[0][1])  # the larger one again, after the smaller
    for (name, s), (e, g) in zip(cases, static):
        with lib.Context(s) as ctx:
            e_c, g_c = ctx.energy_grad()
        d = np.linalg.norm(g - g_c) / np.linalg.norm(g_c)
        print(f"{name}: static {e:.12f} context {e_c:.12f} oracle {CONFIGS[name]['energy']:.12f} |dg|/|g| {d:.2e}")
        assert abs(e - e_c) < 1e-10
        assert d <= 1e-10
        assert abs(e - CONFIGS[name]["energy"]) < 1e-9
    assert abs(again[0] - static[0][0]) < 1e-10
    assert np.linalg.norm(again[1] - static[0][1]) <= 1e-10 * np.linalg.norm(static[0][1])


def test_static_entry_refuses_what_it_cannot_hold(lib):
    with pytest.raises(lib.VsvbError, match="MiB of device scratch"):
        lib.energy_grad_static(workloads.config4())   # LiF128: 1536 AOs
    with pytest.raises(lib.VsvbError, match="npair = 0"):
        lib.energy_grad_static(workloads.config2())   # spin-coupled
