"""The Fortran host path on the GPU: tests/fortran/vsvb_drive (built by
__graft_entry__.build() from valence_amd/fortran/vsvb_hip.F90 and
tests/fortran/vsvb_drive.F90) fills vsvb_system_t from Fortran arrays, calls
vsvb_context_create / vsvb_energy / vsvb_energy_grad through bind(C)
interfaces, and must reproduce the oracle's energy (1e-9 Eh), the ctypes
path's gradient (1e-10 relative) and central differences of the oracle
energy (1e-7 relative)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from valence_amd import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "fortran", "_build", "vsvb_drive")
CONFIGS = json.load(open(os.path.join(GOLDEN, "config_energies.json")))
GRAD_FD = json.load(open(os.path.join(GOLDEN, "grad_fd.json")))


@pytest.mark.parametrize("which", ["config0", "config2"])
def test_fortran_driver_energy_and_gradient(tmp_path, which):
    sys.path.insert(0, os.path.join(HERE, "fortran"))
    import dump
    from valence_amd import lib as L
    assert os.path.exists(EXE), f"{EXE} missing: run __graft_entry__.build() (tests/fortran/build_driver.py)"
    s = workloads.config0() if which == "config0" else workloads.config2()
    path = str(tmp_path / f"{which}.sys")
    dump.write(s, path)
    print(f"fortran driver: {which}", flush=True)
    r = subprocess.run([EXE, path], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    vals = {ln.split()[0]: float(ln.split()[1]) for ln in r.stdout.splitlines() if len(ln.split()) == 2}
    g = np.fromfile(path + ".grad", dtype=np.float64)
    assert g.shape == (s.totlen,)
    ref = CONFIGS[which]
    assert abs(vals["energy"] - ref["energy"]) < 1e-9
    assert abs(vals["energy_grad"] - vals["energy"]) < 1e-10
    L.load()
    with L.Context(s) as ctx:
        _, g_c = ctx.energy_grad()
    assert np.linalg.norm(g - g_c) <= 1e-10 * np.linalg.norm(g_c)
    d = GRAD_FD[which]["fd"]
    idx = np.array(sorted(int(k) for k in d))
    fd = np.array([d[str(k)] for k in idx])
    rel = np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)
    print(f"{which}: E={vals['energy']:.12f} oracle={ref['energy']:.12f} |g - fd|/|fd|={rel:.2e}")
    assert rel < 1e-7
