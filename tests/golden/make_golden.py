"""Write tests/golden/config_energies.json: oracle energies of the seeded workloads.

For each workload the oracle (the C restatement of the reference algorithm) is
run twice: with the reference's own thresholds (itol = 10^-ntol_i, the
orbital-quartet Schwarz and post-evaluation screens of valence.F90:1189-1190,
1286-1287) and with every screen disabled (itol = 0), so the effect of those
screens on the energy is recorded next to the value.
Run: python tests/golden/make_golden.py [config index or extra workload name ...]
"""
import json
import os
import sys
import time

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from valence_amd import workloads  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "config_energies.json")


def main(which):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for i in which:
        s = workloads.EXTRA[i]() if i in workloads.EXTRA else workloads.CONFIGS[i]()
        t = time.time()
        ref = oracle.energy(s, nthreads=os.cpu_count())
        t_ref = time.time() - t
        t = time.time()
        uns = oracle.energy(s, itol=0.0, nthreads=os.cpu_count())
        t_uns = time.time() - t
        key = i if i in workloads.EXTRA else f"config{i}"
        data[key] = dict(title=s.title, energy=ref.energy, energy_unscreened=uns.energy,
                                  screen_effect=ref.energy - uns.energy, e_nuc=ref.e_nuc, ntasks=ref.ntasks,
                                  n_schwarz_skip=ref.n_schwarz_skip, n_post_skip=ref.n_post_skip,
                                  n_int2e_task=ref.n_int2e_task, n_pairs=s.nspatial * (s.nspatial + 1) // 2,
                                  n_dets=ref.n_dets,
                                  seconds=t_ref, seconds_unscreened=t_uns, threads=os.cpu_count())
        print(f"{key}: E={ref.energy:.12f} unscreened={uns.energy:.12f} diff={ref.energy - uns.energy:.3e} "
              f"({t_ref:.1f}s, {t_uns:.1f}s)", flush=True)
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main([x if x in workloads.EXTRA else int(x) for x in sys.argv[1:]] or [0, 1, 2])
