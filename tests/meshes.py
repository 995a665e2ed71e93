"""Shared synthetic inputs for the tests."""
import numpy as np

import pierce_amd as pz


def cube_mesh(centres, edges):
    """Axis-aligned cubes (the reference's 12-triangle cube, generate_test_cubes.py:19-33)."""
    tmpl = pz.generate(pz.GEN_CUBES, count=1, seed=0, radius_min=2.0, radius_max=2.0, box=0.0)
    unit = tmpl.vertices  # edge 2, centred at the origin: half extents 1
    verts, idx, t2o = [], [], []
    for o, (c, e) in enumerate(zip(centres, edges)):
        base = len(verts)
        for v in unit:
            verts.append(np.float32(c) + np.float32(e / 2.0) * v)
        idx.extend((tmpl.indices + base).tolist())
        t2o.extend([o] * len(tmpl.indices))
    m = pz.Mesh(np.array(verts, np.float32), np.array(idx, np.uint32), np.array(t2o, np.int32))
    return pz.extract_edges(m)


def equal_cube_offsets():
    import itertools
    return [np.array(d, np.float32) for d in itertools.product([-2.5, -2.0, -1.0, 0.0, 0.5, 2.0, 2.25], repeat=3)]


def nested_cubes(seed=8, n=30):
    import random
    rng = random.Random(seed)
    cent, edges = [], []
    for _ in range(2 * n):
        cent.append(np.array([rng.choice([0, 0.5, 1, 1.5, 3]) for _ in range(3)], np.float32))
        edges.append(rng.choice([1.0, 2.0, 4.0, 6.0]))
    return cube_mesh(cent[:n], edges[:n]), cube_mesh(cent[n:], edges[n:])


def containment_sets(n=300, box=60.0):
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=7, level_min=2, level_max=2, radius_min=0.5, radius_max=2.0,
                    box=box)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=n, seed=8, level_min=2, level_max=2, radius_min=0.5, radius_max=2.0,
                    box=box, copy_fraction=0.2, copy_scale_min=0.3, copy_scale_max=0.8, copy_source_count=n,
                    copy_source_seed=7)
    return a, b

