"""Build the HIP library in-tree: valence_amd/libvsvb_hip.so (gfx950).

Steps: regenerate the ERI class code (csrc/gen_eri.py), compile every
translation unit with hipcc in parallel, link one shared library.  Objects go
to valence_amd/_build/.  ``build(force=True)`` recompiles everything.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvsvb_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["vsvb_core.hip", "vsvb_sc.hip", "vsvb_eri_coop.hip"] + [f"vsvb_eri_b{b}.hip" for b in range(6)]
ERI_HEADERS = ["vsvb_device.h", "vsvb_eri_kernels.h", "eri_gen.inc", "coop_tables.inc"]
HOST_HEADERS = ["vsvb_internal.h", "eri_flops.inc", os.path.join("..", "..", "include", "vsvb_hip.h")]

# Packed FP32 instructions are disabled for all device code (they have given
# silent wrong results on this hardware pool); the host pass ignores the flag.
CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(HERE, "..", "include")]


def _digest(src: str) -> str:
    h = hashlib.sha256()
    for f in [src] + (ERI_HEADERS + HOST_HEADERS if not src.startswith("vsvb_eri_b") else ERI_HEADERS):
        p = os.path.join(CSRC, f)
        if os.path.exists(p):
            with open(p, "rb") as fh:
                h.update(fh.read())
    h.update(" ".join(CFLAGS).encode())
    return h.hexdigest()


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    stamp = obj + ".sha"
    dig = _digest(src)
    if not force and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == dig:
        return obj
    cmd = [HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    with open(stamp, "w") as fh:
        fh.write(dig)
    return obj


def build(force: bool = False, jobs: int = 0) -> str:
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_eri.py")], check=True, stdout=subprocess.DEVNULL)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        tmp = LIB + ".tmp"
        r = subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr[-4000:])
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
