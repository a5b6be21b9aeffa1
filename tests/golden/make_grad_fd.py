"""Write tests/golden/grad_fd.json: central differences of the oracle energy.

For each case, dE/dc_k = (E(c + h e_k) - E(c - h e_k)) / 2h with h = 1e-5
(the step the gradient gate names), E being the oracle's energy (the C
restatement of the reference algorithm, reference screens on).  The GPU
gradient tests compare the adjoint gradient with these values at 1e-7
norm-wise relative.

Cases and weights:
* config0 (water): every weight;
* config2 (H12 spin-coupled) and lif8 (LiF 2x2x2): 24 weights drawn with
  numpy.random.default_rng(4200 + case index);
* the reference's spin-coupled inputs testing/be-sc, testing/h2o-vdz-sc1,
  examples/lih.SCval and its NDF inputs testing/be+ndf, testing/h+ndf,
  testing/cu+, examples/cu+.3d10 (orbitals expanded in NDFs, NDF weights
  included): every weight, at the seeded perturbation the GPU tests use
  (their recorded weights are optimized, |dE/dc| ~ 1e-5):
  c + 0.05 * N(0,1)[default_rng(7)] * max(|c|, 0.1).
For the NDF inputs also a directional derivative along a seeded unit
direction u (numpy.random.default_rng(3300 + case index), over all weights)
by the five-point stencil with h = 1e-3, whose rounding error (~1e-15 |E| / h)
stays below 1e-7 of the derivative also where the gradient is small next to
|E| (examples/cu+.3d10: |g| ~ 0.2 at |E| ~ 1600 Eh, where the h = 1e-5
differences carry ~3e-7 of rounding noise).
Run: python tests/golden/make_grad_fd.py [nthreads]
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from valence_amd import workloads  # noqa: E402
from valence_amd.system import read_input  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "grad_fd.json")
H = 1e-5
RECORDED_SC = ["testing/be-sc", "testing/h2o-vdz-sc1", "examples/lih.SCval"]
RECORDED_NDF = ["testing/be+ndf", "testing/h+ndf", "testing/cu+", "examples/cu+.3d10"]


def perturbed(s):
    """The seeded perturbation of an optimized recorded input (same as the GPU tests)."""
    rng = np.random.default_rng(7)
    s.coeff = s.coeff + 0.05 * rng.standard_normal(s.totlen) * np.maximum(np.abs(s.coeff), 0.1)
    return s


def case_system(name):
    if name == "config0":
        return workloads.config0()
    if name == "config2":
        return workloads.config2()
    if name == "lif8":
        return workloads.lif8()
    grp, base = name.split("/")
    return perturbed(read_input(os.path.join(ROOT, "tests", "golden", "inputs", f"{grp}__{base}")))


def case_weights(name, s):
    if name in ("config2", "lif8"):
        seed = 4200 + ["config2", "lif8"].index(name)
        return sorted(np.random.default_rng(seed).choice(s.totlen, 24, replace=False).tolist())
    return list(range(s.totlen))


def main(nthreads):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    data["h"] = H
    for name in ["config0", "config2", "lif8"] + RECORDED_SC + RECORDED_NDF:
        s = case_system(name)
        ent = data.setdefault(name, {"title": s.title, "fd": {}})
        for k in case_weights(name, s):
            if str(k) in ent["fd"]:
                continue
            sp, sm = s.copy(), s.copy()
            sp.coeff[k] += H
            sm.coeff[k] -= H
            ep = oracle.energy(sp, nthreads=nthreads).energy
            em = oracle.energy(sm, nthreads=nthreads).energy
            ent["fd"][str(k)] = (ep - em) / (2 * H)
            json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)
        print(f"{name}: {len(ent['fd'])} components", flush=True)
    for ci, name in enumerate(RECORDED_NDF):
        ent = data[name]
        if "dir5" in ent:
            continue
        s = case_system(name)
        u = np.random.default_rng(3300 + ci).standard_normal(s.totlen)
        u /= np.linalg.norm(u)
        hd = 1e-3
        ev = {}
        for m in (-2, -1, 1, 2):
            t = s.copy()
            t.coeff = s.coeff + m * hd * u
            ev[m] = oracle.energy(t, nthreads=nthreads).energy
        ent["dir5"] = {"seed": 3300 + ci, "h": hd,
                       "value": (-ev[2] + 8 * ev[1] - 8 * ev[-1] + ev[-2]) / (12 * hd)}
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)
        print(f"{name}: directional derivative {ent['dir5']['value']:.12e}", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1))
