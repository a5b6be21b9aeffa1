"""Compile the Fortran binding (valence_amd/fortran/vsvb_hip.F90) and its test
driver (tests/fortran/vsvb_drive.F90) with amdflang, linked against
valence_amd/libvsvb_hip.so (rpath relative to the binary, so the tree can move).
Output: tests/fortran/_build/vsvb_drive.  Run: python tests/fortran/build_driver.py"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "vsvb_drive")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def build() -> str:
    os.makedirs(OUT, exist_ok=True)
    mod_src = os.path.join(ROOT, "valence_amd", "fortran", "vsvb_hip.F90")
    drv_src = os.path.join(HERE, "vsvb_drive.F90")
    libdir = os.path.join(ROOT, "valence_amd")
    steps = [
        [FLANG, "-O2", "-cpp", "-module-dir", OUT, "-c", mod_src, "-o", os.path.join(OUT, "vsvb_hip.o")],
        [FLANG, "-O2", "-cpp", "-I" + OUT, "-module-dir", OUT, "-c", drv_src, "-o", os.path.join(OUT, "vsvb_drive.o")],
        [FLANG, "-o", EXE, os.path.join(OUT, "vsvb_drive.o"), os.path.join(OUT, "vsvb_hip.o"), "-L" + libdir,
         "-lvsvb_hip", "-Wl,-rpath,$ORIGIN/../../../valence_amd"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(" ".join(cmd) + "\n" + r.stdout[-3000:] + r.stderr[-3000:])
    return EXE


if __name__ == "__main__":
    print(build())
    sys.exit(0)
