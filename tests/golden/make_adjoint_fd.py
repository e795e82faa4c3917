"""Write tests/golden/adjoint_fd_257x257x50.json: the adjoint workload's
objective and its central finite differences at 16 seeded initial theta and
16 seeded initial u values, from the HIP dycore (plain FP64).

Workload: the 64x64x40 warm-bubble setup (flat, uniform eta, neutral 300 K,
+2 K bubble at 2 km, u = 10 m/s, 0.01 K noise, dt 6 s, ns 6) on 257 x 257 x 50;
J = sum(w**2) over w levels f >= 25 after 50 large steps; eps = 1e-6 (K for
theta, m/s for u); seeds 5 (theta points) and 6 (u points).

Run on a GPU:  python tests/golden/make_adjoint_fd.py OUT.json
"""

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from wrf_gpu_amd import ic, sensitivity  # noqa: E402

N_STEPS = 50
EPS = 1e-6


def main(path):
    case = ic.bubble_3d(nx=257, ny=257, nz=50)
    t0 = time.time()
    j0 = sensitivity.objective(case, N_STEPS)
    print(f"J = {j0!r} ({time.time() - t0:.1f}s)", flush=True)
    out = {"workload": "warm bubble 257x257x50, dx 1 km, dt 6 s, ns 6", "n_steps": N_STEPS, "eps": EPS,
           "objective": "sum(w**2) over w levels f >= (nz+1)//2", "J": j0, "points": []}
    for field, seed in (("theta", 5), ("u", 6)):
        pts = sensitivity.seeded_points(case, field, 16, seed)
        for r in sensitivity.central_differences(case, N_STEPS, field, pts, EPS):
            out["points"].append({"field": r.field, "index": list(r.index), "j_plus": r.j_plus,
                                  "j_minus": r.j_minus, "derivative": r.derivative})
            print(r.field, r.index, r.derivative, flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1])
