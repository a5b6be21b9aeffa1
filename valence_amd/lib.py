"""ctypes binding of libvsvb_hip.so (the C ABI in include/vsvb_hip.h).

This is the in-process replacement of the reference's process boundary for
energies (vtools runs the ``valence`` binary through a pipe,
vtools/iotools.py:487-495).  The library must have been built in-tree
(``python -m valence_amd.build``); loading fails loudly if it is missing --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from .system import VsvbSystem, pack

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvsvb_hip.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

VSVB_OK = 0


class VsvbError(RuntimeError):
    pass


class CSystem(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("natom", "ntype", "nshell_type", "nprim", "norbs", "mxctr", "totlen",
                                          "npair", "nunpd", "ndocc", "ndf", "nspinc", "ntol_d", "ntol_i")] + [
        ("atom_type", _i32p), ("coords", _f64p), ("nuc_charge", _f64p),
        ("type_shell_start", _i32p), ("shell_l", _i32p), ("shell_prim_start", _i32p),
        ("prim_exp", _f64p), ("prim_coef", _f64p),
        ("orb_natom", _i32p), ("orb_atoms", _i32p), ("orb_start", _i32p), ("xpset", _i32p),
        ("coeff", _f64p), ("pair_sc", _i32p), ("coeff_sc", _f64p)]


class COptions(C.Structure):
    _fields_ = [("quartet_tau", C.c_double), ("device", C.c_int32), ("pair_chunk", C.c_int32),
                ("grid_inverse_min", C.c_int32), ("beta_only_last", C.c_int32)]


EXPORTS = ["vsvb_context_create", "vsvb_context_destroy", "vsvb_energy", "vsvb_energy_parts", "vsvb_energy_grad",
           "vsvb_energy_grad_device", "vsvb_batch_create", "vsvb_batch_run", "vsvb_batch_destroy",
           "vsvb_last_error", "vsvb_last_counters", "vsvb_set_profiling", "vsvb_class_stats",
           "vsvb_batch_set_profiling", "vsvb_batch_class_stats", "vsvb_debug_buffer", "vsvb_debug_boys",
           "vsvb_version", "vsvb_batch_uses_graph", "vsvb_batch_tuning", "vsvb_batch_set_graph",
           "vsvb_batch_set_lanes", "vsvb_batch_set_group_launch", "vsvb_batch_launches", "vsvb_debug_gemm",
           "vsvb_spin_matrices", "vsvb_spin_opt", "vsvb_set_spin_coupling", "vsvb_surface",
           "vsvb_batch_set_coords", "vsvb_surface_batch", "vsvb_energy_grad_static"]

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise VsvbError(f"{path} not found: build it with `python -m valence_amd.build` (no CPU fallback exists)")
    L = C.CDLL(path)
    L.vsvb_context_create.argtypes = [C.POINTER(CSystem), C.POINTER(COptions), C.POINTER(C.c_void_p)]
    L.vsvb_context_destroy.argtypes = [C.c_void_p]
    L.vsvb_context_destroy.restype = None
    L.vsvb_energy.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
    L.vsvb_energy_parts.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    L.vsvb_energy_grad.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
    L.vsvb_energy_grad_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.vsvb_batch_create.argtypes = [C.c_int32, C.POINTER(C.POINTER(CSystem)), C.POINTER(COptions),
                                    C.POINTER(C.c_void_p)]
    L.vsvb_batch_run.argtypes = [C.c_void_p, C.c_int32, _f64p, C.POINTER(_f64p)]
    L.vsvb_batch_destroy.argtypes = [C.c_void_p]
    L.vsvb_batch_destroy.restype = None
    L.vsvb_last_error.restype = C.c_char_p
    L.vsvb_last_counters.argtypes = [C.c_void_p, _f64p]
    L.vsvb_debug_buffer.argtypes = [C.c_void_p, C.c_char_p, _f64p, C.c_int64]
    L.vsvb_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    L.vsvb_debug_boys.argtypes = [C.c_int32, _f64p, C.c_int32, _f64p]
    L.vsvb_surface.argtypes = [C.c_void_p, _f64p, C.POINTER(C.c_double)]
    L.vsvb_batch_set_coords.argtypes = [C.c_void_p, _f64p]
    L.vsvb_surface_batch.argtypes = [C.c_void_p, _f64p, _f64p]
    L.vsvb_spin_matrices.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    L.vsvb_spin_opt.argtypes = [C.c_void_p, _f64p, C.POINTER(C.c_double), _f64p]
    L.vsvb_set_spin_coupling.argtypes = [C.c_void_p, _f64p]
    L.vsvb_debug_gemm.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p,
                                  C.c_double, _f64p]
    L.vsvb_class_stats.argtypes = [C.c_void_p, _f64p]
    L.vsvb_batch_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    L.vsvb_batch_class_stats.argtypes = [C.c_void_p, _f64p]
    L.vsvb_version.restype = C.c_char_p
    L.vsvb_batch_uses_graph.argtypes = [C.c_void_p]
    L.vsvb_batch_tuning.argtypes = [C.c_void_p, _f64p]
    L.vsvb_batch_set_graph.argtypes = [C.c_void_p, C.c_int32]
    L.vsvb_batch_set_lanes.argtypes = [C.c_void_p, C.c_int32]
    L.vsvb_batch_set_group_launch.argtypes = [C.c_void_p, C.c_int32]
    L.vsvb_batch_launches.argtypes = [C.c_void_p, _f64p, C.c_int32, C.POINTER(C.c_int32)]
    L.vsvb_energy_grad_static.argtypes = [C.POINTER(CSystem), C.POINTER(COptions), _f64p, _f64p]
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray], t):
    return None if a is None else a.ctypes.data_as(t)


def _check(rc: int, what: str):
    if rc != VSVB_OK:
        msg = load().vsvb_last_error().decode(errors="replace")
        raise VsvbError(f"{what} failed ({rc}): {msg}")


def device_boys(L: int, T, compact: bool = False) -> np.ndarray:
    """F_0..F_L at the given arguments, evaluated on the device from the full table
    (cooperative and one-electron kernels) or, compact=True, from the per-L table
    the straight-line quartet kernels stage in LDS."""
    T = np.ascontiguousarray(T, np.float64).ravel()
    out = np.zeros((len(T), L + 1))
    _check(load().vsvb_debug_boys(L + (16 if compact else 0), _ptr(T, _f64p), len(T), _ptr(out, _f64p)),
           "vsvb_debug_boys")
    return out


def device_gemm(A: np.ndarray, B: np.ndarray, ta: bool = False, tb: bool = False, alpha: float = 1.0,
                beta: float = 0.0, C0: Optional[np.ndarray] = None) -> np.ndarray:
    """alpha op(A) op(B) + beta C0 through the library's device GEMM (tests)."""
    A = np.ascontiguousarray(A, np.float64); B = np.ascontiguousarray(B, np.float64)
    M, K = (A.shape[1], A.shape[0]) if ta else A.shape
    N = B.shape[0] if tb else B.shape[1]
    C = np.zeros((M, N)) if C0 is None else np.array(C0, np.float64, copy=True, order="C")
    _check(load().vsvb_debug_gemm(M, N, K, int(ta), int(tb), alpha, _ptr(A, _f64p), _ptr(B, _f64p), beta,
                                  _ptr(C, _f64p)), "vsvb_debug_gemm")
    return C


class _Packed:
    def __init__(self, s: VsvbSystem):
        pk = pack(s)
        # the ABI takes pair_sc in the reference's storage order, pair_sc(npair, 2, nspinc)
        pk.i32["pair_sc"] = np.ascontiguousarray(np.asarray(pk.i32["pair_sc"], np.int32).transpose(0, 2, 1))
        self.keep = {**pk.i32, **pk.f64}
        sc = pk.scalars
        self.cs = CSystem(**{k: sc[k] for k in ("natom", "ntype", "nshell_type", "nprim", "norbs", "mxctr", "totlen",
                                                "npair", "nunpd", "ndocc", "ndf", "nspinc", "ntol_d", "ntol_i")},
                          **{k: _ptr(v, _i32p) for k, v in pk.i32.items()},
                          **{k: _ptr(v, _f64p) for k, v in pk.f64.items()})


def energy_grad_static(s: VsvbSystem, device: int = 0, want_grad: bool = True):
    """Energy and gradient through the allocation-free entry (vsvb_energy_grad_static):
    no context, every device buffer in the library's static scratch array."""
    L = load()
    pk = _Packed(s)
    opt = COptions(device=device)
    e = C.c_double()
    g = np.zeros(s.totlen) if want_grad else None
    _check(L.vsvb_energy_grad_static(C.byref(pk.cs), C.byref(opt), C.byref(e), _ptr(g, _f64p)),
           "vsvb_energy_grad_static")
    return e.value, g


class Context:
    """One molecule/basis/orbital layout resident on the GPU (vsvb_context)."""

    def __init__(self, s: VsvbSystem, quartet_tau: float = 0.0, device: int = 0, pair_chunk: int = 0,
                 grid_inverse_min: int = 0, beta_only_last: bool = False):
        L = load()
        self.sys = s
        self._packed = _Packed(s)
        opt = COptions(quartet_tau=quartet_tau, device=device, pair_chunk=pair_chunk,
                       grid_inverse_min=grid_inverse_min, beta_only_last=int(beta_only_last))
        h = C.c_void_p()
        _check(L.vsvb_context_create(C.byref(self._packed.cs), C.byref(opt), C.byref(h)), "vsvb_context_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().vsvb_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def energy(self, coeff: Optional[np.ndarray] = None, coords: Optional[np.ndarray] = None,
               parts: bool = False):
        cf = None if coeff is None else np.ascontiguousarray(coeff, np.float64)
        xy = None if coords is None else np.ascontiguousarray(coords, np.float64).reshape(-1)
        e = C.c_double()
        p = np.zeros(4)
        _check(load().vsvb_energy(self._h, _ptr(xy, _f64p), _ptr(cf, _f64p), C.byref(e), _ptr(p, _f64p)),
               "vsvb_energy")
        return (e.value, p) if parts else e.value

    def surface(self, x_angstrom: np.ndarray) -> float:
        """calcsurface: energy in cm^-1 at Cartesian coordinates in angstrom (vsvb_surface)."""
        x = np.ascontiguousarray(x_angstrom, np.float64).reshape(-1)
        v = C.c_double()
        _check(load().vsvb_surface(self._h, _ptr(x, _f64p), C.byref(v)), "vsvb_surface")
        return v.value

    def spin_matrices(self, coeff: Optional[np.ndarray] = None):
        """(ham, ovl), [nspinc, nspinc] each: E = c.ham.c / c.ovl.c + E_nuc (vsvb_spin_matrices)."""
        m = int(self.sys.nspinc)
        h = np.zeros((m, m)); o = np.zeros((m, m))
        cf = None if coeff is None else np.ascontiguousarray(coeff, np.float64)
        _check(load().vsvb_spin_matrices(self._h, _ptr(cf, _f64p), _ptr(h, _f64p), _ptr(o, _f64p)),
               "vsvb_spin_matrices")
        return h, o

    def spin_opt(self, coeff: Optional[np.ndarray] = None):
        """(energy, coeff_sc) of the lowest spin-coupling root (vsvb_spin_opt)."""
        m = int(self.sys.nspinc)
        c = np.zeros(m); e = C.c_double()
        cf = None if coeff is None else np.ascontiguousarray(coeff, np.float64)
        _check(load().vsvb_spin_opt(self._h, _ptr(cf, _f64p), C.byref(e), _ptr(c, _f64p)), "vsvb_spin_opt")
        return e.value, c

    def set_spin_coupling(self, coeff_sc: np.ndarray):
        c = np.ascontiguousarray(coeff_sc, np.float64)
        _check(load().vsvb_set_spin_coupling(self._h, _ptr(c, _f64p)), "vsvb_set_spin_coupling")

    def energy_grad(self, coeff: Optional[np.ndarray] = None, coords: Optional[np.ndarray] = None):
        cf = None if coeff is None else np.ascontiguousarray(coeff, np.float64)
        xy = None if coords is None else np.ascontiguousarray(coords, np.float64).reshape(-1)
        e = C.c_double()
        g = np.zeros(self.sys.totlen)
        _check(load().vsvb_energy_grad(self._h, _ptr(xy, _f64p), _ptr(cf, _f64p), C.byref(e), _ptr(g, _f64p)),
               "vsvb_energy_grad")
        return e.value, g

    def energy_grad_device(self, coeff_ptr: int, energy_ptr: int, grad_ptr: int, stream: int = 0) -> None:
        """E and dE/dcoeff with device pointers (e.g. torch CUDA tensors' data_ptr()),
        ordered on the given HIP stream handle (vsvb_energy_grad_device)."""
        _check(load().vsvb_energy_grad_device(self._h, C.c_void_p(coeff_ptr), C.c_void_p(energy_ptr),
                                              C.c_void_p(grad_ptr), C.c_void_p(stream or None)),
               "vsvb_energy_grad_device")

    def buffer(self, name: str, n: int) -> np.ndarray:
        """Copy of an internal device buffer (diagnostics)."""
        out = np.zeros(n)
        _check(load().vsvb_debug_buffer(self._h, name.encode(), _ptr(out, _f64p), n), "vsvb_debug_buffer")
        return out

    def set_profiling(self, on: bool = True):
        _check(load().vsvb_set_profiling(self._h, int(on)), "vsvb_set_profiling")

    def class_stats(self) -> np.ndarray:
        """[21, 6]: ms, quartets, primitive quartets, skipped, algorithmic FP64 ops, launches (last call)."""
        out = np.zeros((21, 6))
        _check(load().vsvb_class_stats(self._h, _ptr(out, _f64p)), "vsvb_class_stats")
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros(4)
        _check(load().vsvb_last_counters(self._h, _ptr(out, _f64p)), "vsvb_last_counters")
        return out


class Batch:
    """Many wave functions of uneven size evaluated in one graph launch (vsvb_batch)."""

    def __init__(self, systems: Sequence[VsvbSystem], quartet_tau: float = 0.0, device: int = 0,
                 pair_chunk: int = 0, grid_inverse_min: int = 0):
        L = load()
        self.systems = list(systems)
        self._packed = [_Packed(s) for s in self.systems]
        arr = (C.POINTER(CSystem) * len(self._packed))(*[C.pointer(p.cs) for p in self._packed])
        self._arr = arr
        opt = COptions(quartet_tau=quartet_tau, device=device, pair_chunk=pair_chunk,
                       grid_inverse_min=grid_inverse_min)
        h = C.c_void_p()
        _check(L.vsvb_batch_create(len(self.systems), arr, C.byref(opt), C.byref(h)), "vsvb_batch_create")
        self._h = h

    def run(self, want_grad: bool = True):
        n = len(self.systems)
        e = np.zeros(n)
        grads: List[np.ndarray] = [np.zeros(s.totlen) for s in self.systems]
        gp = (_f64p * n)(*[_ptr(g, _f64p) for g in grads]) if want_grad else None
        _check(load().vsvb_batch_run(self._h, int(want_grad), _ptr(e, _f64p), gp), "vsvb_batch_run")
        return e, (grads if want_grad else None)

    def set_profiling(self, on: bool = True):
        _check(load().vsvb_batch_set_profiling(self._h, int(on)), "vsvb_batch_set_profiling")

    def set_graph(self, on: bool = True):
        """Off: every run is enqueued directly (no graph capture or replay)."""
        _check(load().vsvb_batch_set_graph(self._h, int(on)), "vsvb_batch_set_graph")

    def set_lanes(self, n: int):
        """Number of streams the groups run on (1 serializes them)."""
        _check(load().vsvb_batch_set_lanes(self._h, int(n)), "vsvb_batch_set_lanes")

    def set_group_launch(self, chunks: int):
        """Class launches over the groups of each stream, in this many chunks per
        stream (0: one pass per group)."""
        _check(load().vsvb_batch_set_group_launch(self._h, int(chunks)), "vsvb_batch_set_group_launch")

    def launches(self) -> np.ndarray:
        """The last run's group launches: rows of (class, kernel: 0 straight-line,
        1 transposed, 2 cooperative, groups, stream, ms, FLOPs)."""
        n = C.c_int32(0)
        _check(load().vsvb_batch_launches(self._h, None, 0, C.byref(n)), "vsvb_batch_launches")
        out = np.zeros((max(n.value, 1), 6))
        _check(load().vsvb_batch_launches(self._h, _ptr(out, _f64p), n.value, C.byref(n)), "vsvb_batch_launches")
        return out[:n.value]

    def set_coords(self, coords_bohr):
        """New coordinates for every wave function (list of [natom, 3] arrays, bohr)."""
        x = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).reshape(-1) for c in coords_bohr]))
        _check(load().vsvb_batch_set_coords(self._h, _ptr(x, _f64p)), "vsvb_batch_set_coords")

    def surface(self, x_angstrom) -> np.ndarray:
        """Energies in cm^-1 of every wave function at new coordinates in angstrom (vsvb_surface_batch)."""
        x = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).reshape(-1) for c in x_angstrom]))
        v = np.zeros(len(self.systems))
        _check(load().vsvb_surface_batch(self._h, _ptr(x, _f64p), _ptr(v, _f64p)), "vsvb_surface_batch")
        return v

    def uses_graph(self) -> bool:
        """True once runs replay the captured HIP graph (captured after the first run)."""
        r = load().vsvb_batch_uses_graph(self._h)
        if r < 0:
            _check(r, "vsvb_batch_uses_graph")
        return bool(r)

    def tuning(self) -> np.ndarray:
        """[2, 21, 8] launch choices of the batch tuner (see vsvb_batch_tuning)."""
        out = np.zeros((2, 21, 8))
        _check(load().vsvb_batch_tuning(self._h, _ptr(out, _f64p)), "vsvb_batch_tuning")
        return out

    def graph_error(self) -> str:
        load().vsvb_batch_uses_graph(self._h)
        return load().vsvb_last_error().decode(errors="replace")

    def class_stats(self) -> np.ndarray:
        out = np.zeros((21, 6))
        _check(load().vsvb_batch_class_stats(self._h, _ptr(out, _f64p)), "vsvb_batch_class_stats")
        return out

    def close(self):
        if getattr(self, "_h", None):
            load().vsvb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
