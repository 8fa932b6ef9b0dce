"""Writes the exact-recovery fixtures from the reference's own data files
(paper_resources/random_12345_data/*.data: whitespace-separated rows, the
last column the noise-free target; random_12345_truth/all.f: the generating
formula; random_12345_results/random_noise_0.txt: the reference's reported
result).  Only data is copied: the first 1000 rows of each dataset, and the
formula and reported expression as text (never evaluated).

    python tests/golden/make_recovery.py /root/reference
"""
import json
import os
import sys

import numpy as np

NAMES = ["random_04_02_0010000_02", "random_04_01_0010000_00", "random_05_02_0010000_04",
         "random_04_02_0010000_04", "random_05_02_0010000_02", "random_06_02_0010000_04",
         "random_07_03_0010000_01"]
ROWS = 1000


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    base = os.path.join(ref, "paper_resources")
    truth = {}
    with open(os.path.join(base, "random_12345_truth", "all.f")) as f:
        for line in f:
            parts = line.rstrip("\n").split("\t")
            if len(parts) == 2:
                truth[parts[0]] = parts[1]
    reported = {}
    with open(os.path.join(base, "random_12345_results", "random_noise_0.txt")) as f:
        for line in f:
            fields = line.rstrip("\n").split("\t")
            kv = dict(x.split("=", 1) for x in fields[1:] if "=" in x)
            reported[fields[0].replace(".data", "")] = {"R2": kv.get("R2"), "exprSimp": kv.get("exprSimp")}
    meta = {}
    for name in NAMES:
        d = np.loadtxt(os.path.join(base, "random_12345_data", name + ".data"))[:ROWS]
        cols = ["x%d" % i for i in range(d.shape[1] - 1)] + ["y"]
        np.savetxt(os.path.join(here, "recovery_%s.csv" % name), d, delimiter=",", header=",".join(cols),
                   comments="", fmt="%.17g")
        meta[name] = {"truth": truth.get(name), "reference_result": reported.get(name), "rows": ROWS}
    with open(os.path.join(here, "recovery.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
