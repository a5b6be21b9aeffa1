"""Write tests/golden/spin_matrices.json: the spin-coupling matrices of the
recorded spin-coupled inputs with more than one spin coupling, from the oracle
(the C restatement of the reference algorithm, every screen off).

The energy numerator (E1 + E2) and norm of a wave function are quadratic in its
spin-coupling weights c: num(c) = c.H.c and N(c) = c.S.c.  So one oracle run per
unit vector e_I gives H_II and S_II, and one per pair e_I + e_J gives
H_IJ = (num(e_I + e_J) - H_II - H_JJ) / 2 (same for S).
Run: python tests/golden/make_spin_matrices.py
"""
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from valence_amd.system import read_input  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "spin_matrices.json")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")


def parts(s, c):
    r = oracle.energy(dataclasses.replace(s, coeff_sc=np.asarray(c, float)), itol=0.0, nthreads=os.cpu_count())
    return r.e1 + r.e2, r.wfnorm, r.e_nuc


def main(names):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        s = read_input(os.path.join(INPUTS, name))
        m = int(s.nspinc)
        t = time.time()
        H = np.zeros((m, m)); S = np.zeros((m, m)); enuc = 0.0
        for i in range(m):
            e = np.zeros(m); e[i] = 1.0
            H[i, i], S[i, i], enuc = parts(s, e)
        for i in range(m):
            for j in range(i):
                e = np.zeros(m); e[i] = e[j] = 1.0
                h, o, _ = parts(s, e)
                H[i, j] = H[j, i] = 0.5 * (h - H[i, i] - H[j, j])
                S[i, j] = S[j, i] = 0.5 * (o - S[i, i] - S[j, j])
        c = np.asarray(s.coeff_sc, float)
        data[name] = dict(nspinc=m, ham=H.tolist(), ovl=S.tolist(), e_nuc=enuc, coeff_sc=c.tolist(),
                          energy_at_coeff_sc=float(c @ H @ c / (c @ S @ c) + enuc), seconds=time.time() - t)
        print(f"{name}: nspinc={m} E(coeff_sc)={data[name]['energy_at_coeff_sc']:.12f} ({time.time() - t:.1f}s)",
              flush=True)
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["testing__be-sc", "testing__be-scv3s+2sc", "examples__be.2SC", "examples__be2s3s.2SC",
                          "testing__n2.sc4val-b.p2"])
