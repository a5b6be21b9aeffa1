"""Write tests/golden/config3_fixtures.json: oracle values for the 96-wave-function batch.

The batch (valence_amd.workloads.config3, seed 1003) is 48 displaced C8H18
geometries, 32 Givens-rotated C8H18 wave functions and 16 alkanes CnH2n+2.
This script runs the oracle (the C restatement of the reference algorithm,
with the reference's own screens at itol = 10^-ntol_i) on:

* energies: displaced members 0-3 and every alkane member (80-95);
* gradients by central differences of the oracle energy (h = 1e-5, the
  step of the gradient gate):
  - member 85 (C3H8): 32 seeded weight components;
  - members 85, 48 (rotated C8H18) and 0 (displaced C8H18): directional
    derivatives along seeded random unit vectors d over all weights,
    (E(c + h d) - E(c - h d)) / 2h, which a gradient g must match as g . d.

The 32 rotated members need no oracle run: a rotation of two doubly occupied
orbitals leaves the energy unchanged, so their energy is config 1's
(tests/golden/config_energies.json).  Directions are regenerated from their
seeds with numpy.random.default_rng(seed).standard_normal(totlen), normalized.
Results are written after every piece, so an interrupted run resumes.
Run: python tests/golden/make_config3.py [nthreads]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from valence_amd import workloads  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "config3_fixtures.json")
H = 1e-5
ENERGY_MEMBERS = [0, 1, 2, 3] + list(range(80, 96))
COMPONENT_MEMBER, N_COMPONENTS = 85, 32
DIRECTION_MEMBERS = {85: [3101, 3102], 48: [3103, 3104], 0: [3105]}


def direction(seed: int, n: int) -> np.ndarray:
    d = np.random.default_rng(seed).standard_normal(n)
    return d / np.linalg.norm(d)


def main(nthreads: int):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    data.setdefault("h", H)
    data.setdefault("energies", {})
    data.setdefault("components", {})
    data.setdefault("directions", {})
    members = workloads.config3()

    def save():
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)

    def E(s):
        return oracle.energy(s, nthreads=nthreads).energy

    for i in ENERGY_MEMBERS:
        if str(i) in data["energies"]:
            continue
        t = time.time()
        e = E(members[i])
        data["energies"][str(i)] = dict(title=members[i].title, energy=e, seconds=time.time() - t)
        print(f"member {i} {members[i].title}: E={e:.12f} ({time.time() - t:.0f}s)", flush=True)
        save()

    s = members[COMPONENT_MEMBER]
    ks = sorted(np.random.default_rng(3100).choice(s.totlen, N_COMPONENTS, replace=False).tolist())
    comp = data["components"].setdefault(str(COMPONENT_MEMBER), {"title": s.title, "fd": {}})
    for k in ks:
        if str(k) in comp["fd"]:
            continue
        sp, sm = s.copy(), s.copy()
        sp.coeff[k] += H
        sm.coeff[k] -= H
        comp["fd"][str(k)] = (E(sp) - E(sm)) / (2 * H)
        save()
    print(f"member {COMPONENT_MEMBER}: {len(comp['fd'])} components", flush=True)

    for i, seeds in DIRECTION_MEMBERS.items():
        s = members[i]
        ent = data["directions"].setdefault(str(i), {"title": s.title, "fd": {}})
        for seed in seeds:
            if str(seed) in ent["fd"]:
                continue
            t = time.time()
            d = direction(seed, s.totlen)
            sp, sm = s.copy(), s.copy()
            sp.coeff = sp.coeff + H * d
            sm.coeff = sm.coeff - H * d
            ent["fd"][str(seed)] = (E(sp) - E(sm)) / (2 * H)
            print(f"member {i} direction {seed}: {ent['fd'][str(seed)]:.10e} ({time.time() - t:.0f}s)", flush=True)
            save()


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1))
