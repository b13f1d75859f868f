"""Extracts the reference's own recorded outputs for the two-sphere
experiment into a small fixture (tests/golden/exp_two_spheres.npz).

demo/exp.py (reference tree) computes V_01 of two 1000-disk Fibonacci sphere
clouds (create_sphere_pointcloud, unit radius), the second translated by
(d, 0, 0) for d = 0, 0.1, ..., 2.0, on the Lebedev 302 x L=32 k-grid, with the
reference's C++ CUDA and C++ CPU (TBB) paths, and records the volumes in
demo/exp_results.csv.  This script copies, for each distance, the volume of
the first run of each method and the analytic lens volume (gt_volume).

Run (needs the reference tree):
  python tests/golden/make_exp_vectors.py /root/reference/demo/exp_results.csv
"""
import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(path):
    rows = list(csv.DictReader(open(path)))
    out = {}
    for method, key in (("C++ CPU", "v_cpu"), ("C++ CUDA", "v_cuda")):
        sel = sorted((float(r["distance"]), float(r["volume"]), float(r["gt_volume"]))
                     for r in rows if r["method"] == method and r["run"] == "1")
        out["distance"] = np.array([s[0] for s in sel])
        out[key] = np.array([s[1] for s in sel])
        out["v_analytic"] = np.array([s[2] for s in sel])
    np.savez(os.path.join(HERE, "exp_two_spheres.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/demo/exp_results.csv")
