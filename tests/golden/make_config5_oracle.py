"""Writes config5_oracle.json: every request of config5_requests.json solved
by the CPU oracle (ebSolve driven by oracle/bsam_oracle.py with the host
layer's setup, tagging and level construction), as compact records: boxes
per level, V-cycles, residual histories (exact doubles), final relative
residuals and per-level SHA-256 digests of Q (tests/config5.py).

Run from the repository root:  python tests/golden/make_config5_oracle.py [workers]
"""
import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
for p in (os.path.join(ROOT, "bsam-in-matlab_amd"), ROOT, os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)


def _one(i):
    import warnings

    warnings.filterwarnings("ignore")
    import config5

    return config5.oracle_record(config5.fixture()["requests"][i])


def main():
    import config5

    n = len(config5.fixture()["requests"])
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    with Pool(workers) as pool:
        recs = pool.map(_one, range(n), chunksize=4)
    with open(os.path.join(HERE, "config5_oracle.json"), "w") as f:
        json.dump({"requests": recs}, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
