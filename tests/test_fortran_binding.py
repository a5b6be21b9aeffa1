"""The Fortran side of the boundary, compiled: the ISO_C_BINDING module
valence_amd/fortran/vsvb_hip.F90 and a driver program are built with amdflang
and linked against libvsvb_hip.so, and the bind(C) types must lay out exactly
as include/vsvb_hip.h does under the C compiler (offsets from a gcc probe).
Running that program on the GPU is tests/test_gpu_fortran.py."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, ".."))

PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "vsvb_hip.h"
int main(void) {
  printf("sizeof vsvb_system %zu\n", sizeof(vsvb_system));
  printf("offset coords %zu\n", offsetof(vsvb_system, coords));
  printf("offset xpset %zu\n", offsetof(vsvb_system, xpset));
  printf("offset coeff_sc %zu\n", offsetof(vsvb_system, coeff_sc));
  printf("sizeof vsvb_options %zu\n", sizeof(vsvb_options));
  printf("offset grid_inverse_min %zu\n", offsetof(vsvb_options, grid_inverse_min));
  printf("offset beta_only_last %zu\n", offsetof(vsvb_options, beta_only_last));
  return 0;
}
"""


def _layout(lines):
    out = {}
    for ln in lines.splitlines():
        parts = ln.split()
        if len(parts) == 3 and parts[0] in ("sizeof", "offset"):
            out[(parts[0], parts[1])] = int(parts[2])
    return out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/amdflang"), reason="amdflang not in this image")
def test_fortran_binding_builds_and_matches_c_layout(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(HERE, "fortran"))
    import build_driver
    exe = build_driver.build()
    f = subprocess.run([exe, "--layout"], capture_output=True, text=True, timeout=60)
    assert f.returncode == 0, f.stderr
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    cc = shutil.which("gcc") or "cc"
    subprocess.run([cc, "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "probe")], check=True)
    c = subprocess.run([str(tmp_path / "probe")], capture_output=True, text=True, check=True)
    lf, lc = _layout(f.stdout), _layout(c.stdout)
    print("Fortran:", lf)
    print("C:      ", lc)
    assert len(lc) == 7 and lf == lc
