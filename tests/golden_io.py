"""Loading the golden fixtures of tests/golden (written by tests/golden/make_golden.py)."""
import os

import numpy as np

import pierce_amd as pz

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("ellipsoids", "tubes", "mix")
PREDICATES = {"overlap": pz.PRED_OVERLAP, "intersection": pz.PRED_INTERSECTION, "containment": pz.PRED_CONTAINMENT}


def meshes(case):
    return (pz.read_pre(os.path.join(GOLDEN, f"{case}_mesh1.pre")),
            pz.read_pre(os.path.join(GOLDEN, f"{case}_mesh2.pre")))


def expected(case, pred):
    path = os.path.join(GOLDEN, f"{case}_{pred}.csv")
    with open(path) as f:
        assert f.readline().strip() == "object_id_mesh1,object_id_mesh2"
        rows = [tuple(int(x) for x in line.split(",")) for line in f if line.strip()]
    return np.array(rows, np.int32).reshape(-1, 2)
