"""Writes the committed golden fixtures of tests/golden/ (run in a container
that has g++ and the reference tree at /root/reference):

  stdlib_random.json   libstdc++ std::mt19937 / uniform_real_distribution vectors
  nn_linear_l2.npz     reference NearestNeighborsLinear on 2000 R^7 points
  nn_linear_se3.npz    reference NearestNeighborsLinear on 2000 SE3 states

The NN fixtures come from oracle/_ref/ref_linear, which compiles the
reference's datastructures/NearestNeighborsLinear.h unchanged.  Only data
(inputs and outputs) is committed.
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle_py  # noqa: E402


def stdlib_random():
    with tempfile.TemporaryDirectory() as td:
        exe = Path(td) / "stdlib_random"
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(HERE / "stdlib_random.cpp")], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    (HERE / "stdlib_random.json").write_text(json.dumps(data, indent=1) + "\n")


def random_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def nn_fixture(name, metric, data, queries, k, r):
    oracle_py.build()
    with tempfile.TemporaryDirectory() as td:
        knn_res, rad_res, nearest_res = oracle_py.ref_linear(metric, data, queries, k, r, Path(td))
    kidx = np.stack([x[0] for x in knn_res]).astype(np.int64)
    kdist = np.stack([x[1] for x in knn_res])
    roff = np.zeros(len(rad_res) + 1, np.int64)
    roff[1:] = np.cumsum([len(x[0]) for x in rad_res])
    ridx = np.concatenate([x[0] for x in rad_res]).astype(np.int64)
    rdist = np.concatenate([x[1] for x in rad_res])
    np.savez_compressed(HERE / name, data=data, queries=queries, k=k, radius=r, knn_idx=kidx, knn_dist=kdist,
                        rad_off=roff, rad_idx=ridx, rad_dist=rdist, nearest=np.array(nearest_res, np.int64))


def main():
    stdlib_random()
    rng = np.random.default_rng(2024)
    data = rng.uniform(-np.pi, np.pi, (2000, 7)).astype(np.float32)
    queries = rng.uniform(-np.pi, np.pi, (64, 7)).astype(np.float32)
    nn_fixture("nn_linear_l2.npz", 0, data, queries, 10, 2.5)
    t = rng.uniform(-1, 1, (2000, 3))
    se3 = np.concatenate([t, random_quaternions(rng, 2000)], axis=1).astype(np.float32)
    tq = rng.uniform(-1, 1, (64, 3))
    se3q = np.concatenate([tq, random_quaternions(rng, 64)], axis=1).astype(np.float32)
    nn_fixture("nn_linear_se3.npz", 1, se3, se3q, 16, 0.6)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
