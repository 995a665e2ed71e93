#!/usr/bin/env python3
"""Writes the golden fixtures of this directory.

Inputs are small seeded generated sets, stored as .pre v4 files so the
fixtures do not depend on the generator.  Expected outputs come from brute
force over all object pairs, never from a BVH:
- overlap(a, b): some triangle of a meets some triangle of b, by exact
  all-pairs triangle-triangle tests (oracle/oracle.cpp or_bruteforce_overlap);
- containment(a, b): not overlap and the first vertex of b's first triangle
  is strictly inside a, by exact +Z ray parity over all of a's triangles;
- intersection(a, b): overlap, or either first vertex inside the other
  object.

These are the semantics of pierce_overlap, pierce_containment and
pierce_intersection (pierce/query/src/applications/pierce_containment_main.cpp:152-169,
pierce_intersection_main.cpp:103-155).  CSVs are sorted by (mesh1, mesh2) and
use the reference's overlap header.

Run from the repository root: python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import pierce_amd as pz  # noqa: E402
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    """name -> (mesh1, mesh2) of each fixture."""
    ell = dict(level_min=1, level_max=1, radius_min=1.0, radius_max=3.0, box=12.0, jitter=0.02)
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=16, seed=101, **ell)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=16, seed=102, copy_fraction=0.4, copy_scale_min=0.3,
                    copy_scale_max=0.8, copy_source_count=16, copy_source_seed=101, **ell)
    tube = dict(box=20.0, radius_min=0.3, radius_max=0.8, tri_median=400, tri_min=200, tri_max=800, tri_sigma=0.3,
                step=1.0)
    ta = pz.generate(pz.GEN_TUBES, count=4, seed=201, **tube)
    tb = pz.generate(pz.GEN_TUBES, count=4, seed=202, degenerate_fraction=0.75, copy_source_count=4,
                     copy_source_seed=201, **tube)
    mix = dict(box=14.0, tri_min=200, tri_median=300, tri_max=600, level_min=1, level_max=2)
    ma = pz.generate(pz.GEN_QUERY_MIX, count=10, seed=301, **mix)
    mb = pz.generate(pz.GEN_QUERY_MIX, count=10, seed=302, **mix)
    return {"ellipsoids": (a, b), "tubes": (ta, tb), "mix": (ma, mb)}


def first_vertex(m: pz.Mesh, obj: int) -> np.ndarray:
    t = int(np.nonzero(m.triangle_to_object == obj)[0][0])
    return m.vertices[m.indices[t, 0]]


def brute_force(m1: pz.Mesh, m2: pz.Mesh):
    ov = {tuple(p) for p in orc.bruteforce_overlap(m1, m2).tolist()}
    n1, n2 = m1.num_objects, m2.num_objects
    inside_12 = np.zeros((n1, n2), bool)  # first vertex of b inside a
    inside_21 = np.zeros((n1, n2), bool)  # first vertex of a inside b
    for i in range(n1):
        for j in range(n2):
            inside_12[i, j] = orc.point_in_object(m1, i, first_vertex(m2, j))
            inside_21[i, j] = orc.point_in_object(m2, j, first_vertex(m1, i))
    pairs = [(i, j) for i in range(n1) for j in range(n2)]
    overlap = sorted(ov)
    containment = [p for p in pairs if p not in ov and inside_12[p]]
    intersection = [p for p in pairs if p in ov or inside_12[p] or inside_21[p]]
    return {"overlap": overlap, "containment": containment, "intersection": intersection}


def write_csv(path, pairs):
    with open(path, "w") as f:
        f.write("object_id_mesh1,object_id_mesh2\n")
        for a, b in pairs:
            f.write(f"{a},{b}\n")


def main():
    for name, (m1, m2) in cases().items():
        pz.write_pre(os.path.join(HERE, f"{name}_mesh1.pre"), m1)
        pz.write_pre(os.path.join(HERE, f"{name}_mesh2.pre"), m2)
        for pred, pairs in brute_force(m1, m2).items():
            write_csv(os.path.join(HERE, f"{name}_{pred}.csv"), pairs)
            print(f"{name} {pred}: {len(pairs)} pairs ({m1.num_objects} x {m2.num_objects} objects, "
                  f"{m1.indices.shape[0]} + {m2.indices.shape[0]} triangles)")


if __name__ == "__main__":
    main()
