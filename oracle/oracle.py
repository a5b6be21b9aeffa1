"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` load this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_givens.so")

_i32p = C.POINTER(C.c_int)
_f64p = C.POINTER(C.c_double)


class OrcSystem(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("natom", "ntype", "nshell_type", "nprim", "norbs", "mxctr", "totlen",
                                        "npair", "nunpd", "ndocc", "ndf", "nspinc", "ntol_d", "ntol_i")] + [
        ("atom_type", _i32p), ("coords", _f64p), ("nuc_charge", _f64p),
        ("type_shell_start", _i32p), ("shell_l", _i32p), ("shell_prim_start", _i32p),
        ("prim_exp", _f64p), ("prim_coef", _f64p),
        ("orb_natom", _i32p), ("orb_atoms", _i32p), ("orb_start", _i32p), ("xpset", _i32p),
        ("coeff", _f64p), ("pair_sc", _i32p), ("coeff_sc", _f64p)]


class OrcOptions(C.Structure):
    _fields_ = [("itol", C.c_double), ("dtol", C.c_double), ("nthreads", C.c_int), ("eri_table", C.c_int),
                ("task_begin", C.c_longlong), ("task_end", C.c_longlong), ("skip_1e", C.c_int), ("plant", C.c_int)]


class OrcResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("e_nuc", C.c_double), ("e1", C.c_double), ("e2", C.c_double),
                ("wfnorm", C.c_double), ("ntasks", C.c_longlong), ("tasks_done", C.c_longlong),
                ("n_schwarz_skip", C.c_longlong), ("n_post_skip", C.c_longlong), ("n_dets", C.c_longlong),
                ("n_int2e_task", C.c_longlong), ("n_int2e_schwarz", C.c_longlong),
                ("seconds_setup", C.c_double), ("seconds_1e", C.c_double), ("seconds_2e", C.c_double)]


_lib = None


def build(force: bool = False) -> None:
    """Compile the oracle (gcc) in place."""
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("vsvb_oracle.c", "vsvb_oracle.h")):
        subprocess.run(["make", "-C", HERE, "liborc.so"] + (["-B"] if force else []), check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_energy.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcOptions), C.POINTER(OrcResult)]
        L.orc_energy.restype = C.c_int
        L.orc_orbital_integrals.argtypes = [C.POINTER(OrcSystem), C.c_double, _f64p, _f64p, _f64p]
        L.orc_normalized_coeff.argtypes = [C.POINTER(OrcSystem), _f64p]
        L.orc_boys.argtypes = [C.c_int, C.c_double, _f64p]
        L.orc_norm_prim.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.orc_givens_single.argtypes = [_f64p, C.c_int, C.c_int, C.c_double]
        L.orc_givens_det.argtypes = [_f64p, C.c_int, C.c_int, C.c_double]
        L.orc_givens_det.restype = C.c_double
        L.orc_eri_block.argtypes = ([_f64p, C.c_int, C.c_int, _f64p, _f64p] * 4) + [_f64p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class _Keep:
    """Holds the numpy buffers an OrcSystem points into."""

    def __init__(self, s):
        from valence_amd.system import pack
        pk = pack(s)
        self.arrays = {**pk.i32, **pk.f64}
        sc = pk.scalars
        self.cs = OrcSystem(
            natom=sc["natom"], ntype=sc["ntype"], nshell_type=sc["nshell_type"], nprim=sc["nprim"],
            norbs=sc["norbs"], mxctr=sc["mxctr"], totlen=sc["totlen"], npair=sc["npair"], nunpd=sc["nunpd"],
            ndocc=sc["ndocc"], ndf=sc["ndf"], nspinc=sc["nspinc"], ntol_d=sc["ntol_d"], ntol_i=sc["ntol_i"],
            **{k: _ptr(v, _i32p) for k, v in pk.i32.items()}, **{k: _ptr(v, _f64p) for k, v in pk.f64.items()})


@dataclass
class OracleEnergy:
    energy: float
    e_nuc: float
    e1: float
    e2: float
    wfnorm: float
    ntasks: int
    tasks_done: int
    n_schwarz_skip: int
    n_post_skip: int
    n_dets: int
    n_int2e_task: int
    n_int2e_schwarz: int
    seconds_setup: float
    seconds_1e: float
    seconds_2e: float


def energy(s, itol: Optional[float] = None, dtol: Optional[float] = None, nthreads: int = 0,
           eri_table: bool = True, task_begin: int = 0, task_end: int = -1, skip_1e: bool = False,
           plant: int = 0) -> OracleEnergy:
    """Plain VSVB energy of ``s`` (a VsvbSystem) by the restated reference algorithm."""
    keep = _Keep(s)
    opt = OrcOptions(itol=-1.0 if itol is None else itol, dtol=-1.0 if dtol is None else dtol,
                     nthreads=nthreads, eri_table=int(eri_table), task_begin=task_begin, task_end=task_end,
                     skip_1e=int(skip_1e), plant=plant)
    res = OrcResult()
    rc = lib().orc_energy(C.byref(keep.cs), C.byref(opt), C.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle failed ({rc})")
    return OracleEnergy(**{f[0]: getattr(res, f[0]) for f in OrcResult._fields_})


def orbital_integrals(s, dtol: Optional[float] = None, with_eri: bool = True):
    """Spatial-orbital S, h and (pq|rs) (n^4) after the reference's normalization."""
    keep = _Keep(s)
    n = s.nspatial
    S = np.zeros((n, n)); h = np.zeros((n, n))
    g = np.zeros((n, n, n, n)) if with_eri else None
    dt = 10.0 ** -s.ntol_d if dtol is None else dtol
    rc = lib().orc_orbital_integrals(C.byref(keep.cs), dt, _ptr(S, _f64p), _ptr(h, _f64p),
                                     _ptr(g, _f64p) if g is not None else None)
    if rc != 0:
        raise RuntimeError("oracle failed")
    return S, h, g


def normalized_coeff(s) -> np.ndarray:
    keep = _Keep(s)
    out = np.zeros(s.totlen)
    lib().orc_normalized_coeff(C.byref(keep.cs), _ptr(out, _f64p))
    return out


def boys(nmax: int, T: float) -> np.ndarray:
    F = np.zeros(nmax + 1)
    lib().orc_boys(nmax, T, _ptr(F, _f64p))
    return F


def eri_block(shells) -> np.ndarray:
    """shells: 4 tuples (center[3], l, exps, normalized coefs) -> raw block (i slowest)."""
    args = []
    keep = []
    size = 1
    for (A, l, e, c) in shells:
        A = np.ascontiguousarray(A, np.float64); e = np.ascontiguousarray(e, np.float64)
        c = np.ascontiguousarray(c, np.float64)
        keep += [A, e, c]
        args += [_ptr(A, _f64p), l, len(e), _ptr(e, _f64p), _ptr(c, _f64p)]
        size *= (l + 1) * (l + 2) // 2
    out = np.zeros(size)
    lib().orc_eri_block(*args, _ptr(out, _f64p))
    return out


def norm_prim(l: int, exps, unnorm) -> np.ndarray:
    e = np.ascontiguousarray(exps, np.float64); u = np.ascontiguousarray(unnorm, np.float64)
    out = np.zeros(len(e))
    lib().orc_norm_prim(l, len(e), _ptr(e, _f64p), _ptr(u, _f64p), _ptr(out, _f64p))
    return out


def givens_det(a: np.ndarray, tol: float) -> float:
    """Determinant by the restated givens_single; ``a`` is an (n, n) matrix."""
    af = np.asfortranarray(a, np.float64)
    return lib().orc_givens_det(_ptr(af, _f64p), a.shape[0], a.shape[0], tol)


def ref_givens_det(a: np.ndarray, tol: float, variant: str = "fortran") -> float:
    """Same determinant through the reference's own compiled kernels (oracle/_ref)."""
    L = C.CDLL(REF_LIB_PATH)
    af = np.array(a, dtype=np.float64, order="F", copy=True)
    n = C.c_int(a.shape[0]); t = C.c_double(tol)
    if variant == "fortran":
        L.givens_single_(_ptr(af, _f64p), C.byref(n), C.byref(n), C.byref(t))
    else:
        L.givensc_(_ptr(af, _f64p), C.byref(n), C.byref(n), C.byref(t))
    return float(np.prod(np.diag(af)))
