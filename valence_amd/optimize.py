"""Orbital optimization driven by the adjoint gradient: the consumer of
vsvb_energy_grad that replaces the reference optimizer's energy differences.

The reference's minimize_energy (src/valence.F90:2744-2885) optimizes the
orbitals of each optimization group (the begin/end orbital labels of the
control line, ``orbset``) one orbital at a time, by a first-order H/S
eigenproblem (first_order_opt, :527-844) or per-weight perturbations
(demgs_opt, :351-521), repeating until the energy changes by less than
10^-ntol kcal/mol for ntol = ntol_e_min..ntol_e_max, with a spin-coupling
optimization (spin_opt, :850-936) after each orbital sweep when nspinc > 1.

Two methods:

* ``minimize_energy`` (the gradient consumer): the weights of every orbital of
  the groups are optimized together by L-BFGS (scipy) on the exact energy
  gradient, one library call per energy and gradient; the other orbitals stay
  fixed, as in the reference.  With nspinc > 1 each orbital pass is followed by
  vsvb_spin_opt and the new coupling weights, until the energy changes by less
  than the finest tolerance.  It converges to the variational minimum over the
  group weights, which is what the reference's first-order method reaches.
* ``dem_minimize``: the reference's direct energy minimization (demgs_opt,
  chosen when ptbnmax > 0 and there are no excitations, :200), restated step
  by step on the library's energies: per weight, +perturb, then -perturb, each
  followed by the orbital's renormalization (normal, :2157-2182), kept if the
  energy drops, otherwise the weight is restored and locked; sweeps repeat
  until no weight moves, then perturb *= feather.  With feather = 0 (the
  reference's DEM test inputs) it is a single-step coordinate search that stops
  above the variational minimum, so its converged energy is the reference's
  own figure, not the minimum (testing/testing.py:151: li_opt).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import lib
from .system import VsvbSystem, nuclear_repulsion, write_orbitals

KCAL_PER_HARTREE = 627.509469  # the reference's tokcal (src/valence.F90:2759)


@dataclass
class OptResult:
    energy: float
    coeff: np.ndarray
    coeff_sc: np.ndarray
    iterations: int
    gradient_norm: float            # over the optimized weights, at the end
    history: List[float] = field(default_factory=list)
    # the energy of the returned weights; differs from ``energy`` after a DEM
    # whose last trials failed (see dem_minimize)
    energy_at_weights: float = float("nan")


# ---------------------------------------------------------------------------
# First-order orbital step (first_order_opt, src/valence.F90:527-844)
# ---------------------------------------------------------------------------
# The reference replaces orbital iorb, in the bra and in the ket, by each
# function of its optimization space in turn (an AO, or an NDF where its
# weight refers to one), evaluates <Psi_i|H|Psi_j> and <Psi_i|Psi_j> for every
# pair (ham, ovl; :674-704), solves ham w = E ovl w (rsg) and takes the lowest
# root, or root(i) + 1 for an excited orbital xorb(i) (:790-805), as the new
# weights; the eigenvalue plus E_nuc is the new energy.  A doubly occupied
# orbital is made singly occupied for this: only its alpha copy is replaced
# (:566-590), and with unpaired electrons the beta-replaced matrices are added
# (spin averaging, :706-748).
#
# The wave function is linear in the replaced orbital's weights x, so its
# energy is the Rayleigh quotient E(x) = E_nuc + x^T H x / x^T S x of the
# pencil (H, S) = (ham, ovl) up to a common factor, and the library's exact
# gradient satisfies, at any x,
#     H x - E_el(x) S x - (x^T S x) grad_x E / 2 = 0,
# n equations that are linear in the unknown (H, S).  Evaluated at n + 4
# points they fix (H, S) up to that common factor (the null vector of the
# stacked equations), which is all the eigenproblem needs.  A doubly occupied
# orbital is evaluated on a copy of the system in which the replacement is a
# separate orbital holding the alpha electron and the original keeps only the
# beta one (vsvb_options.beta_only_last).  Spin-averaged pencils are each
# scaled to x^T S x = 1 at the current weights before they are added: both
# determinants then have the same norm there, so the sum matches the
# reference's up to one common factor.


def _pencil_from_gradients(eval_fn, c: np.ndarray, seed: int = 0):
    """(H, S), symmetric, with c^T S c = 1, of E_el(x) = x^T H x / x^T S x from
    eval_fn(x) -> (E_el, dE/dx) at n + 4 seeded points around c (not at c
    itself, where a spin-coupled expansion may have a singular transition
    overlap).  The solution is the null vector of the stacked equations; it
    must stand out from the next singular value by a factor of 1e5 at least
    (weights nearly spanned by the other orbitals make the equations
    ill-conditioned along those directions, not ambiguous)."""
    n = c.size
    iu, ju = np.triu_indices(n)
    nv = iu.size
    rng = np.random.default_rng(seed)
    scale = np.linalg.norm(c) / np.sqrt(n)
    pts = [c + 0.5 * scale * rng.standard_normal(n) for _ in range(n + 4)]
    rows = []
    for x in pts:
        e, g = eval_fn(x)
        # coefficient of H_ij / S_ij (i <= j) in (H x)_r and in x^T S x
        A_h = np.zeros((n, nv))
        A_h[iu, np.arange(nv)] += x[ju]
        off = iu != ju
        A_h[ju[off], np.arange(nv)[off]] += x[iu[off]]
        quad = np.where(off, 2.0, 1.0) * x[iu] * x[ju]
        A_s = -e * A_h - 0.5 * np.outer(g, quad)
        blk = np.hstack([A_h, A_s])
        blk /= np.linalg.norm(blk, axis=1, keepdims=True) + 1e-300
        rows.append(blk)
    A = np.vstack(rows)
    _, sv, vt = np.linalg.svd(A, full_matrices=False)
    if not sv[-1] <= 1e-5 * sv[-2]:
        raise RuntimeError(f"orbital pencil not determined by the energies and gradients "
                           f"(singular values {sv[-2]:.3e}, {sv[-1]:.3e})")
    z = vt[-1]
    H = np.zeros((n, n)); S = np.zeros((n, n))
    H[iu, ju] = z[:nv]; H[ju, iu] = z[:nv]
    S[iu, ju] = z[nv:]; S[ju, iu] = z[nv:]
    nrm = float(c @ S @ c)
    return H / nrm, S / nrm


def _substituted(s: VsvbSystem, p: int, beta: bool):
    """Copy of ``s`` in which doubly occupied orbital p is split: a replacement
    x (same optimization space) holds the alpha electron and p the beta one
    (beta: the other way round).  Orbital order: spin-coupled and unpaired
    orbitals, then the alpha holder (as the last unpaired orbital), the other
    doubly occupied ones, the beta holder (as the last doubly occupied
    orbital), the NDFs -- the layout vsvb_options.beta_only_last describes.
    Returns the system, assemble(c, x) (its weight vector from ``s``'s weights
    c and x) and the slice of x in it."""
    nnd, nd = 2 * int(s.npair) + int(s.nunpd), int(s.ndocc)
    dl = list(range(nnd, nnd + nd))
    dl.remove(p)
    X = -1  # the replacement
    order = list(range(nnd)) + [p if beta else X] + dl + [X if beta else p] + list(range(nnd + nd, s.norbs))
    t = s.copy()
    natom, atoms, start, xp, src = [], [], [0], [], []
    for q in order:
        o = p if q == X else q
        sl = slice(int(s.orb_start[o]), int(s.orb_start[o + 1]))
        natom.append(int(s.orb_natom[o]))
        atoms.append(s.orb_atoms[o])
        xp.extend(int(v) for v in s.xpset[sl])
        start.append(start[-1] + sl.stop - sl.start)
        src.append((q, sl))
    t.nunpd = int(s.nunpd) + 1
    t.orb_natom = np.array(natom, np.int32)
    t.orb_atoms = np.array(atoms, np.int32)
    t.orb_start = np.array(start, np.int32)
    t.xpset = np.array(xp, np.int32)
    t.coeff = np.zeros(len(xp))
    t.xorb = np.zeros(0, np.int32)
    t.xroot = np.zeros(0, np.int32)
    t.orbset = np.zeros((0, 2), np.int32)

    def assemble(c, x):
        return np.concatenate([x if q == X else c[sl] for q, sl in src])

    xs = next(slice(int(t.orb_start[k]), int(t.orb_start[k + 1])) for k, (q, _) in enumerate(src) if q == X)
    return t, assemble, xs


def first_order_matrices(s: VsvbSystem, c: np.ndarray, iorb: int, device: int = 0, contexts: Optional[dict] = None):
    """(ham, ovl) of first_order_opt for orbital iorb (1-based) at weights c, up
    to a common factor (scaled to c_iorb^T ovl c_iorb = 1 per spin-averaged
    part), and E_nuc.  ``contexts`` caches the library contexts between calls."""
    p = iorb - 1
    nnd = 2 * int(s.npair) + int(s.nunpd)
    nsp = nnd + int(s.ndocc)
    if not 0 <= p < nsp:
        raise ValueError(f"orbital {iorb} is not an occupied orbital")
    sl = slice(int(s.orb_start[p]), int(s.orb_start[p + 1]))
    cache = {} if contexts is None else contexts
    parts = []
    if p < nnd:  # singly occupied (spin-coupled or unpaired): linear in its weights as it is
        key = ("own", p)
        if key not in cache:
            cache[key] = (lib.Context(s, device=device), None, sl)
        variants = [key]
    else:
        variants = [("sub", p, False)] + ([("sub", p, True)] if s.nunpd > 0 else [])
        for key in variants:
            if key not in cache:
                t, assemble, xs = _substituted(s, p, key[2])
                t.coeff = assemble(c, np.array(c[sl]))
                cache[key] = (lib.Context(t, device=device, beta_only_last=True), assemble, xs)
    H = S = None
    enuc = nuclear_repulsion(s)
    for key in variants:
        ctx, assemble, xs = cache[key]

        def eval_fn(x, ctx=ctx, assemble=assemble, xs=xs):
            if assemble is None:
                full = c.copy()
                full[xs] = x
            else:
                full = assemble(c, x)
            e, g = ctx.energy_grad(full)
            return e - enuc, g[xs]

        Hk, Sk = _pencil_from_gradients(eval_fn, np.array(c[sl], dtype=np.float64))
        H = Hk if H is None else H + Hk
        S = Sk if S is None else S + Sk
    return H, S, enuc


def first_order_opt(s: VsvbSystem, c: np.ndarray, iorb: int, device: int = 0,
                    contexts: Optional[dict] = None, norm_ctx=None):
    """One first-order step of orbital iorb (1-based): new weights (normalized,
    normal) in a copy of c and the step's energy (eigenvalue + E_nuc)."""
    from scipy.linalg import eigh
    H, S, enuc = first_order_matrices(s, c, iorb, device, contexts)
    w, V = eigh(H, S)
    root = 0
    for xo, xr in zip(s.xorb, s.xroot):
        if int(xo) == iorb:
            root = int(xr)
    p = iorb - 1
    sl = slice(int(s.orb_start[p]), int(s.orb_start[p + 1]))
    out = np.array(c, dtype=np.float64)
    out[sl] = V[:, root]
    if norm_ctx is not None:  # normal(iorb, iorb): unit self-overlap
        norm_ctx.energy(out)
        Sd = norm_ctx.buffer("S_orb", s.nspatial * s.nspatial).reshape(s.nspatial, s.nspatial)
        out[sl] /= np.sqrt(Sd[p, p])
    return out, float(w[root]) + enuc


def minimize_energy_first_order(s: VsvbSystem, device: int = 0) -> OptResult:
    """The reference's minimize_energy (src/valence.F90:2744-2885) with its
    first-order method (ptbnmax = 0): tolerance sweeps ntol_e_min..ntol_e_max,
    the orbital sets (orbset) optimized orbital by orbital until the energy
    changes by less than 10^-ntol kcal/mol, spin_opt after each sweep when
    nspinc > 1, at most max_iter iterations.  ``energy`` is the energy the
    reference tracks (the last step's eigenvalue + E_nuc); ``energy_at_weights``
    that of the returned weights."""
    c = np.array(s.coeff, dtype=np.float64)
    csc = np.array(s.coeff_sc, dtype=np.float64)
    hist: List[float] = []
    contexts: dict = {}
    ctx = lib.Context(s, device=device)
    try:
        # calculate_vsvb_energy normalizes every orbital before the guess energy (:144-145)
        energy = ctx.energy(c)
        n = s.nspatial
        Sd = ctx.buffer("S_orb", n * n).reshape(n, n)
        for p in range(s.norbs - s.ndf):
            c[int(s.orb_start[p]):int(s.orb_start[p + 1])] /= np.sqrt(Sd[p, p])
        hist.append(energy)
        num_iter = 0
        nset = len(s.orbset)
        for ntol in range(int(s.ntol_e_min), int(s.ntol_e_max) + 1):
            etol = 10.0 ** -ntol
            finished = False
            while not finished:
                eprv_sc = energy
                orbconv = False
                while not orbconv and num_iter < s.max_iter:
                    eprv_orb = energy
                    for b, e in s.orbset:
                        setcnv = False
                        while not setcnv and num_iter < s.max_iter:
                            eprv_set = energy
                            num_iter += 1
                            for iorb in range(int(b), int(e) + 1):
                                c, energy = first_order_opt(s, c, iorb, device, contexts, norm_ctx=ctx)
                            hist.append(energy)
                            setcnv = abs(energy - eprv_set) * KCAL_PER_HARTREE < etol
                    orbconv = abs(energy - eprv_orb) * KCAL_PER_HARTREE < etol
                    if nset == 1:
                        orbconv = True
                if s.nspinc > 1:
                    num_iter += 1
                    energy, csc = ctx.spin_opt(c)
                    ctx.set_spin_coupling(csc)
                    for key in list(contexts):
                        contexts[key][0].set_spin_coupling(csc)
                    hist.append(energy)
                    finished = abs(energy - eprv_sc) * KCAL_PER_HARTREE < etol
                else:
                    finished = True
                if num_iter >= s.max_iter:
                    finished = True
        e_w = ctx.energy(c)
        return OptResult(energy=energy, coeff=c, coeff_sc=csc, iterations=num_iter, gradient_norm=float("nan"),
                         history=hist, energy_at_weights=e_w)
    finally:
        ctx.close()
        for v in contexts.values():
            v[0].close()


def save_orbitals(s: VsvbSystem, r: "OptResult", directory: str, mode: str = "done") -> None:
    """The optimized orbitals as the reference leaves them (xm_output after
    minimize_energy, src/valence.F90:2882): 'orbitals' (and 'nelecwfn' with
    several spin couplings) in ``directory``; mode 'done' adds the tolerance
    10^-ntol_e_max kcal/mol the run converged to."""
    t = s.copy()
    t.coeff = np.asarray(r.coeff, dtype=np.float64).copy()
    t.coeff_sc = np.asarray(r.coeff_sc, dtype=np.float64).copy()
    write_orbitals(t, r.energy, directory, mode, 10.0 ** -int(s.ntol_e_max))


def free_weights(s: VsvbSystem) -> np.ndarray:
    """Indices of the weights of the orbitals in the optimization groups
    (``orbset``, 1-based orbital labels, inclusive ranges)."""
    idx = []
    for b, e in s.orbset:
        for p in range(int(b) - 1, int(e)):
            idx.extend(range(int(s.orb_start[p]), int(s.orb_start[p + 1])))
    return np.array(sorted(set(idx)), dtype=np.int64)


def minimize_energy(s: VsvbSystem, device: int = 0, max_passes: int = 50, gtol: float = 1e-8,
                    etol_kcal: Optional[float] = None) -> OptResult:
    """Minimize the energy of ``s`` over the weights of its optimization groups.
    etol_kcal: convergence of the outer passes (default 10^-ntol_e_max kcal/mol)."""
    from scipy.optimize import minimize

    etol = (10.0 ** -s.ntol_e_max if etol_kcal is None else etol_kcal) / KCAL_PER_HARTREE
    free = free_weights(s)
    if free.size == 0:
        raise ValueError("no orbital optimization groups in the input")
    c = np.array(s.coeff, dtype=np.float64)
    hist: List[float] = []
    iters = 0
    with lib.Context(s, device=device) as ctx:
        e_prev = ctx.energy(c)
        hist.append(e_prev)
        gnorm = np.inf
        for _ in range(max_passes):
            def fun(x):
                cc = c.copy()
                cc[free] = x
                e, g = ctx.energy_grad(cc)
                return e, g[free]

            r = minimize(fun, c[free], jac=True, method="L-BFGS-B",
                         options={"maxiter": 2000, "ftol": 1e-16, "gtol": gtol, "maxcor": 30})
            c[free] = r.x
            iters += int(r.nit)
            # the energy does not depend on an orbital's scale: back to unit norms
            for p in range(s.norbs):
                sl = slice(int(s.orb_start[p]), int(s.orb_start[p + 1]))
                nrm = np.linalg.norm(c[sl])
                if nrm > 0:
                    c[sl] /= nrm
            if s.nspinc > 1:
                e_sc, csc = ctx.spin_opt(c)
                ctx.set_spin_coupling(csc)
            e, g = ctx.energy_grad(c)
            gnorm = float(np.linalg.norm(g[free]))
            hist.append(e)
            if abs(e - e_prev) < etol:
                break
            e_prev = e
        csc_out = np.array(s.coeff_sc, dtype=np.float64)
        if s.nspinc > 1:
            _, csc_out = ctx.spin_opt(c)
        return OptResult(energy=hist[-1], coeff=c, coeff_sc=csc_out, iterations=iters, gradient_norm=gnorm,
                         history=hist, energy_at_weights=hist[-1])


def dem_minimize(s: VsvbSystem, device: int = 0) -> OptResult:
    """The reference's DEM optimizer (demgs_opt inside minimize_energy's loops,
    src/valence.F90:351-521, 2744-2885) on the library's energies.  Returns the
    energy the reference tracks and prints as the total (after a failed trial
    the weight is restored and the previous energy kept, :494-500).  As in the
    reference, that restore resets one weight to its value before the trials,
    while the normalizations after the two trials have rescaled the orbital's
    other weights: the orbital changes shape and the kept energy is no longer
    that of the weights (h2o-vdz-dem: the returned weights lie 5.8e-4 Eh above
    the tracked total).  ``energy_at_weights`` is the energy of the returned
    weights, which is what the 'orbitals' file of either program reproduces."""
    if s.nspinc > 1:
        raise ValueError("DEM with several spin couplings also runs spin_opt; not restated here")
    c = np.array(s.coeff, dtype=np.float64)
    n = s.nspatial
    hist: List[float] = []
    calls = 0
    with lib.Context(s, device=device) as ctx:

        def energy_and_norms(cc):
            nonlocal calls
            calls += 1
            e = ctx.energy(cc)
            S = ctx.buffer("S_orb", n * n).reshape(n, n)
            return e, np.diag(S).copy()

        def normal(p, diag):  # coefficients of orbital p scaled to unit self-overlap
            sl = slice(int(s.orb_start[p]), int(s.orb_start[p + 1]))
            c[sl] *= diag[p] ** -0.5

        # calculate_vsvb_energy normalizes every orbital before the guess energy (:144-145)
        e0, diag = energy_and_norms(c)
        for p in range(s.norbs):
            normal(p, diag)
        energy = e0
        hist.append(energy)

        def demgs_opt(iorb: int, energy: float, etol: float) -> float:
            p = iorb - 1
            lo, hi = int(s.orb_start[p]), int(s.orb_start[p + 1])
            perturb = float(s.ptbnmax)
            while KCAL_PER_HARTREE * perturb ** 2 >= etol:
                lock = [False] * (hi - lo)
                updated = True
                while updated:
                    updated = False
                    for i in range(hi - lo):
                        if lock[i]:
                            continue
                        k = lo + i
                        eprev = energy
                        origc = c[k]
                        c[k] = origc + perturb
                        energy, diag = energy_and_norms(c)
                        normal(p, diag)
                        if energy < eprev:
                            updated = True
                            continue
                        c[k] = origc - perturb
                        energy, diag = energy_and_norms(c)
                        normal(p, diag)
                        if energy < eprev:
                            updated = True
                            continue
                        c[k] = origc
                        lock[i] = True
                        _, diag = energy_and_norms(c)
                        normal(p, diag)
                        energy = eprev
                perturb *= float(s.feather)
            return energy

        num_iter = 0
        nset = len(s.orbset)
        for ntol in range(int(s.ntol_e_min), int(s.ntol_e_max) + 1):
            etol = 10.0 ** -ntol  # kcal/mol, as minimize_energy compares it
            orbconv = False
            while not orbconv and num_iter < s.max_iter:
                eprv_orb = energy
                for b, e in s.orbset:
                    setcnv = False
                    while not setcnv and num_iter < s.max_iter:
                        eprv_set = energy
                        num_iter += 1
                        for iorb in range(int(b), int(e) + 1):
                            energy = demgs_opt(iorb, energy, etol)
                        hist.append(energy)
                        setcnv = abs(energy - eprv_set) * KCAL_PER_HARTREE < etol or int(b) == int(e)
                orbconv = abs(energy - eprv_orb) * KCAL_PER_HARTREE < etol or nset == 1
        e_w = ctx.energy(c)
        return OptResult(energy=energy, coeff=c, coeff_sc=np.array(s.coeff_sc, dtype=np.float64), iterations=calls,
                         gradient_norm=float("nan"), history=hist, energy_at_weights=e_w)
