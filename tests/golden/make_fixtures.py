"""Writes the seeded workload parameters both sides read:

  config2_params.json   variable-D Dirichlet problem (NumPy default_rng(2))
  config5_requests.json served workload: 512 requests (default_rng(5)) and
                        Poisson arrival offsets at 1 request/s (default_rng(55);
                        scale by 1/rate for other rates)

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "bsam-in-matlab_amd"))

from bsamgpu import configs  # noqa: E402


def main():
    with open(os.path.join(HERE, "config2_params.json"), "w") as f:
        json.dump({"seed": 2, **configs.varcoef_params(2)}, f, indent=1)
    reqs = configs.served_params(512, seed=5)
    arr = configs.served_arrivals(512, 1.0, seed=55)
    with open(os.path.join(HERE, "config5_requests.json"), "w") as f:
        json.dump({"seed": 5, "arrival_seed": 55, "arrivals_at_rate_1": arr, "requests": reqs}, f)


if __name__ == "__main__":
    main()
