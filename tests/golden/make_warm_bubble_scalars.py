"""Extract the reference-held warm-bubble snapshot scalars into a fixture.

Reads the reference's committed proof objects (JSON data, loaded with the
json module) and writes only the numbers the oracle tests compare against:
  proofs/v013/cpu_bedrock/f2/skamarock_bubble_diagnostics.json   (JAX-CPU run)
  proofs/v0100/idealized_baseline_u1/warm_bubble_verdict.json     (GPU run)
Run once in a container that has the reference checkout at /root/reference.
"""

import json
import os

REF = "/root/reference"
SOURCES = {
    "jax_cpu_v013": "proofs/v013/cpu_bedrock/f2/skamarock_bubble_diagnostics.json",
    "gpu_v0100": "proofs/v0100/idealized_baseline_u1/warm_bubble_verdict.json",
}
KEYS = ("second", "max_abs_w_m_s", "theta_prime_max_k", "theta_prime_min_k",
        "positive_theta_center_x_m", "positive_theta_center_z_m", "max_abs_u_m_s",
        "mass_total_pa", "theta_symmetry_linf_k")


def main():
    out = {"case": "skamarock warm bubble 2-D slab, 80x40, dx=dz=250 m, dt=0.1 s, 10 sound steps",
           "sources": {}, "snapshots": {}}
    for name, rel in SOURCES.items():
        with open(os.path.join(REF, rel)) as fh:
            data = json.load(fh)
        snaps = data.get("snapshots")
        out["sources"][name] = rel
        out["snapshots"][name] = [{k: s[k] for k in KEYS if k in s} for s in snaps]
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "warm_bubble_2d_reference_scalars.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
