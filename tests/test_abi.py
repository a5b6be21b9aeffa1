"""The C ABI: the built library loads without a GPU and exports every function
include/vsvb_hip.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "vsvb_hip.h")
LIB = os.path.join(ROOT, "valence_amd", "libvsvb_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(vsvb_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("vsvb_context_create", "vsvb_context_destroy", "vsvb_energy", "vsvb_energy_parts",
                 "vsvb_energy_grad", "vsvb_energy_grad_device", "vsvb_batch_create", "vsvb_batch_run",
                 "vsvb_batch_destroy", "vsvb_last_error"):
        assert must in names


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        pytest.skip("libvsvb_hip.so not built (python -m valence_amd.build)")
    return LIB


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_answers_without_a_gpu(built):
    L = ctypes.CDLL(built)
    for n in declared_functions():
        assert hasattr(L, n), n
    L.vsvb_version.restype = ctypes.c_char_p
    assert L.vsvb_version().decode().startswith("vsvb-mi355x")
    L.vsvb_last_error.restype = ctypes.c_char_p
    assert isinstance(L.vsvb_last_error(), bytes)


def test_python_binding_lists_every_symbol():
    from valence_amd import lib
    assert sorted(lib.EXPORTS) == declared_functions()


def test_null_arguments_are_rejected_without_touching_the_device(built):
    L = ctypes.CDLL(built)
    L.vsvb_context_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    assert L.vsvb_context_create(None, None, None) == -1
    L.vsvb_energy.argtypes = [ctypes.c_void_p] * 5
    assert L.vsvb_energy(None, None, None, None, None) == -1
