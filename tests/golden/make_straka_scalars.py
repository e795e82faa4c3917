"""Extract the reference-held Straka density-current snapshot scalars into a fixture.

Reads the reference's committed proof object (JSON data, loaded with the json
module) and writes only the numbers the oracle test compares against:
  proofs/v013/cpu_bedrock/f2/straka_density_current_diagnostics.json  (JAX-CPU run)
the initial snapshot and the 900 s snapshot (9000 steps of dt 0.1 s, constant-K
diffusion with K = 75 m^2/s).  Run once in a container that has the reference
checkout at /root/reference.
"""

import json
import os

REF = "/root/reference"
SOURCE = "proofs/v013/cpu_bedrock/f2/straka_density_current_diagnostics.json"
KEYS = ("second", "max_abs_w_m_s", "max_abs_u_m_s", "theta_prime_min_k", "theta_prime_max_k",
        "cold_theta_center_x_m", "cold_theta_center_z_m", "positive_theta_center_x_m", "positive_theta_center_z_m",
        "front_position_m", "mass_total_pa", "theta_symmetry_linf_k", "w_symmetry_linf_m_s")


def main():
    with open(os.path.join(REF, SOURCE)) as fh:
        data = json.load(fh)
    pick = lambda s: {k: s[k] for k in KEYS if k in s}  # noqa: E731
    out = {"case": "Straka density current 2-D slab, 500x60, dx=dz=100 m, dt=0.1 s, 10 sound steps, K=75 m^2/s",
           "source": SOURCE, "timesteps": data["timesteps"], "dt_s": data["dt_s"],
           "initial": pick(data["initial_snapshot"]), "snapshots": [pick(s) for s in data["snapshots"]]}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "straka_density_current_reference_scalars.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
