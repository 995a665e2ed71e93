"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each check is independent of the oracle's own formulas:
  * orientation signs against exact rational arithmetic (fractions.Fraction on
    the fp32 inputs, which are exact binary fractions);
  * segment/triangle against a parametric + barycentric rational solver (a
    different formulation from the oracle's sign tests);
  * point-in-mesh parity against the convex half-space test on convex meshes,
    with query points placed on vertex/edge projections (the degenerate rays
    simulation of simplicity must resolve);
  * whole joins against closed forms for axis-aligned cubes, brute force on
    tiny sets, and the Face (triangle-triangle) form against the Pierce (edge)
    form and the TOUCH restatement.
"""
import itertools
import random
from fractions import Fraction as F

import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.meshes import cube_mesh, nested_cubes


# ------------------------------------------------------------ exact reference
def fr(v):
    return [F(float(np.float32(x))) for x in v]


def sub(a, b):
    return [x - y for x, y in zip(a, b)]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def sgn(x):
    return (x > 0) - (x < 0)


def det3_rows(r0, r1, r2):
    return dot(r0, cross(r1, r2))


def ref_orient3d(a, b, c, d):
    a, b, c, d = fr(a), fr(b), fr(c), fr(d)
    return sgn(det3_rows(sub(a, d), sub(b, d), sub(c, d)))


def ref_orient2d(a, b, c, i=0, j=1):
    a, b, c = fr(a), fr(b), fr(c)
    return sgn((a[i] - c[i]) * (b[j] - c[j]) - (a[j] - c[j]) * (b[i] - c[i]))


def ref_seg_seg_2d(p, q, a, b):
    """Closed 2-D segments via parametric solve (rationals)."""
    r = [q[0] - p[0], q[1] - p[1]]
    s = [b[0] - a[0], b[1] - a[1]]
    den = r[0] * s[1] - r[1] * s[0]
    ap = [a[0] - p[0], a[1] - p[1]]
    if den != 0:
        t = (ap[0] * s[1] - ap[1] * s[0]) / den
        u = (ap[0] * r[1] - ap[1] * r[0]) / den
        return 0 <= t <= 1 and 0 <= u <= 1
    if ap[0] * r[1] - ap[1] * r[0] != 0:
        return False  # parallel, not collinear
    # collinear: project on the dominant axis of r (or s), check interval overlap
    k = 0 if abs(r[0]) + abs(s[0]) >= abs(r[1]) + abs(s[1]) else 1
    lo1, hi1 = sorted([p[k], q[k]])
    lo2, hi2 = sorted([a[k], b[k]])
    if lo1 == hi1 and lo2 == hi2 and (p != a):
        return False
    return max(lo1, lo2) <= min(hi1, hi2)


def ref_point_in_tri_2d(x, a, b, c):
    def area(u, v, w):
        return (v[0] - u[0]) * (w[1] - u[1]) - (v[1] - u[1]) * (w[0] - u[0])

    T = area(a, b, c)
    l1, l2, l3 = area(x, b, c), area(a, x, c), area(a, b, x)
    if T > 0:
        return l1 >= 0 and l2 >= 0 and l3 >= 0
    return l1 <= 0 and l2 <= 0 and l3 <= 0


def ref_seg_tri(p, q, a, b, c):
    """Closed segment against closed triangle: plane crossing + barycentrics (rationals)."""
    p, q, a, b, c = fr(p), fr(q), fr(a), fr(b), fr(c)
    n = cross(sub(b, a), sub(c, a))
    if n == [0, 0, 0]:
        return False
    dp, dq = dot(n, sub(p, a)), dot(n, sub(q, a))
    if (dp > 0 and dq > 0) or (dp < 0 and dq < 0):
        return False
    if dp == 0 and dq == 0:
        k = max(range(3), key=lambda i: abs(n[i]))
        ii, jj = [m for m in range(3) if m != k]
        P = lambda v: (v[ii], v[jj])
        pp, qq, aa, bb, cc = P(p), P(q), P(a), P(b), P(c)
        if ref_point_in_tri_2d(pp, aa, bb, cc) or ref_point_in_tri_2d(qq, aa, bb, cc):
            return True
        return any(ref_seg_seg_2d(pp, qq, u, v) for u, v in ((aa, bb), (bb, cc), (cc, aa)))
    t = dp / (dp - dq)
    x = [pi + t * (qi - pi) for pi, qi in zip(p, q)]
    nn = dot(n, n)
    u = dot(n, cross(sub(b, x), sub(c, x)))
    v = dot(n, cross(sub(c, x), sub(a, x)))
    w = dot(n, cross(sub(a, x), sub(b, x)))
    assert u + v + w == nn
    return u >= 0 and v >= 0 and w >= 0


# ------------------------------------------------------------ point sets
def degenerate_points(rng, n):
    """Small-integer and shared coordinates: many coplanar / collinear cases."""
    return [np.array([rng.randint(-3, 3) * 0.5 for _ in range(3)], np.float32) for _ in range(n)]


def wide_points(rng, n):
    out = []
    for _ in range(n):
        e = rng.choice([-60, -20, 0, 0, 0, 10, 30])
        out.append(np.array([rng.uniform(-1, 1) * 2.0 ** e for _ in range(3)], np.float32))
    return out


@pytest.mark.parametrize("gen", [degenerate_points, wide_points])
def test_orient3d_matches_rationals(gen):
    rng = random.Random(1)
    pts = gen(rng, 60)
    for _ in range(1500):
        a, b, c, d = rng.sample(pts, 4)
        assert orc.orient3d(a, b, c, d) == ref_orient3d(a, b, c, d)
        assert orc.orient3d_exact(a, b, c, d) == ref_orient3d(a, b, c, d)


def test_orient3d_near_coplanar():
    # d is a float32 rounding of a point on the plane of a, b, c: tiny signed volumes
    rng = random.Random(2)
    for _ in range(500):
        a, b, c = (np.array([rng.uniform(0, 100) for _ in range(3)], np.float32) for _ in range(3))
        s, t = rng.random(), rng.random()
        d = (a.astype(np.float64) + s * (b - a) + t * (c - a)).astype(np.float32)
        assert orc.orient3d(a, b, c, d) == ref_orient3d(a, b, c, d)


def test_orient2d_matches_rationals():
    rng = random.Random(3)
    pts = degenerate_points(rng, 40) + wide_points(rng, 20)
    for _ in range(2000):
        a, b, c = rng.sample(pts, 3)
        i, j = rng.choice([(0, 1), (1, 2), (2, 0)])
        assert orc.orient2d(a, b, c, i, j) == ref_orient2d(a, b, c, i, j)


def test_orient3d_sign_convention():
    # d below the counter-clockwise (seen from +z) triangle -> positive
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    assert orc.orient3d(a, b, c, [0.1, 0.1, -1]) == 1
    assert orc.orient3d(a, b, c, [0.1, 0.1, 1]) == -1
    assert orc.orient3d(a, b, c, [5, 7, 0]) == 0


@pytest.mark.parametrize("gen", [degenerate_points, wide_points])
def test_seg_tri_matches_rational_solver(gen):
    rng = random.Random(4)
    pts = gen(rng, 30)
    hits = 0
    for _ in range(1500):
        p, q, a, b, c = rng.sample(pts, 5)
        exp = ref_seg_tri(p, q, a, b, c)
        hits += exp
        assert orc.seg_tri(p, q, a, b, c) == exp, (p, q, a, b, c)
    if gen is degenerate_points:
        assert hits > 50  # the sample exercises both outcomes


def test_seg_tri_touching_cases():
    a, b, c = [0, 0, 0], [2, 0, 0], [0, 2, 0]
    cases = [
        ([0, 0, -1], [0, 0, 1], True),      # through vertex a
        ([1, 0, -1], [1, 0, 1], True),      # through edge ab
        ([1, 1, -1], [1, 1, 1], True),      # through edge bc (x + y = 2)
        ([1.5, 1.5, -1], [1.5, 1.5, 1], False),
        ([1, 0.5, 0], [1, 0.5, 3], True),   # endpoint on the face
        ([1, 0.5, 1e-7], [1, 0.5, 3], False),
        ([-1, 1, 0], [3, 1, 0], True),      # coplanar crossing
        ([2, 0, 0], [3, 0, 0], True),       # coplanar, touches vertex b
        ([2.5, 0, 0], [3, 0, 0], False),    # coplanar, collinear, disjoint
        ([0.2, 0.2, 0], [0.3, 0.3, 0], True),  # coplanar, inside
    ]
    for p, q, exp in cases:
        assert orc.seg_tri(p, q, a, b, c) == exp, (p, q)
        assert ref_seg_tri(p, q, a, b, c) == exp, (p, q)


def test_tri_tri_symmetric_and_matches_edges():
    rng = random.Random(5)
    pts = degenerate_points(rng, 25)
    for _ in range(800):
        t1 = np.array(rng.sample(pts, 3))
        t2 = np.array(rng.sample(pts, 3))
        r = orc.tri_tri(t1, t2)
        assert r == orc.tri_tri(t2, t1)
        if ref_orient3d(t1[0], t1[1], t1[2], [0, 0, 0]) is not None:
            exp = any(ref_seg_tri(t1[k], t1[(k + 1) % 3], *t2) for k in range(3)) or any(
                ref_seg_tri(t2[k], t2[(k + 1) % 3], *t1) for k in range(3))
            degenerate = (cross(sub(fr(t1[1]), fr(t1[0])), sub(fr(t1[2]), fr(t1[0]))) == [0, 0, 0] or
                          cross(sub(fr(t2[1]), fr(t2[0])), sub(fr(t2[2]), fr(t2[0]))) == [0, 0, 0])
            assert r == (exp and not degenerate)


# ------------------------------------------------------------ point in mesh
def convex_inside(mesh, obj, x):
    """Strictly inside a convex, outward-oriented closed mesh: below every face plane."""
    tris = mesh.indices[mesh.triangle_to_object == obj]
    for t in tris:
        a, b, c = (mesh.vertices[i] for i in t)
        if ref_orient3d(a, b, c, x) <= 0:  # on or above (outside) the face plane
            return False
    return True


def test_point_in_convex_mesh_with_degenerate_rays():
    cube = pz.generate(pz.GEN_CUBES, count=1, seed=3, radius_min=4.0, radius_max=4.0, box=0.0)
    ico = pz.generate(pz.GEN_ELLIPSOIDS, count=1, seed=9, level_min=1, level_max=1, radius_min=2.0, radius_max=2.0,
                      box=0.0)
    rng = random.Random(6)
    for mesh in (cube, ico):
        verts = mesh.vertices
        lo, hi = verts.min(0), verts.max(0)
        pts = []
        for v in verts:  # rays through vertices and along vertical edges/faces
            for dz in (-3.0, -0.5, 0.0, 0.5):
                pts.append(np.array([v[0], v[1], v[2] + dz], np.float32))
        for _ in range(300):
            pts.append(np.array([rng.uniform(lo[k] - 1, hi[k] + 1) for k in range(3)], np.float32))
        for tri in mesh.indices:  # edge midpoints projected: rays grazing edges
            for k in range(3):
                m = (verts[tri[k]].astype(np.float64) + verts[tri[(k + 1) % 3]]) / 2
                pts.append(np.array([m[0], m[1], lo[2] - 1.0], np.float32))
        n_in = 0
        for x in pts:
            on_surface = any(ref_seg_tri(x, x, *(mesh.vertices[i] for i in t)) for t in mesh.indices)
            if on_surface:
                continue
            exp = convex_inside(mesh, 0, x)
            n_in += exp
            assert orc.point_in_object(mesh, 0, x) == exp, x
        assert n_in > 5


# ------------------------------------------------------------ joins
def test_equal_cubes_closed_form():
    # two equal axis-aligned cubes of edge s overlap iff ||d||_inf <= s (face contact included)
    s = 2.0
    offsets = []
    for d in itertools.product([-2.5, -2.0, -1.0, 0.0, 0.5, 2.0, 2.25], repeat=3):
        offsets.append(np.array(d, np.float32))
    A = cube_mesh([np.zeros(3, np.float32)] * len(offsets), [s] * len(offsets))
    B = cube_mesh(offsets, [s] * len(offsets))
    got = orc.face_overlap(A, B)
    expected = set()
    for i in range(len(offsets)):
        for j, d in enumerate(offsets):
            if np.abs(d).max() <= s:
                expected.add((i, j))
    assert set(map(tuple, got.tolist())) == expected
    np.testing.assert_array_equal(orc.edge_overlap(A, B), got)


def test_nested_cubes_closed_form():
    # unequal cubes: overlap iff closed boxes meet and neither is strictly inside the other
    rng = random.Random(8)
    cent, edges = [], []
    for _ in range(60):
        cent.append(np.array([rng.choice([0, 0.5, 1, 1.5, 3]) for _ in range(3)], np.float32))
        edges.append(rng.choice([1.0, 2.0, 4.0, 6.0]))
    A = cube_mesh(cent[:30], edges[:30])
    B = cube_mesh(cent[30:], edges[30:])

    def box(c, e):
        return c - e / 2, c + e / 2

    exp = set()
    for i in range(30):
        la, ha = box(cent[i], edges[i])
        for j in range(30):
            lb, hb = box(cent[30 + j], edges[30 + j])
            meet = np.all(la <= hb) and np.all(lb <= ha)
            a_in_b = np.all(la > lb) and np.all(ha < hb)
            b_in_a = np.all(lb > la) and np.all(hb < ha)
            if meet and not a_in_b and not b_in_a:
                exp.add((i, j))
    got = orc.face_overlap(A, B)
    assert set(map(tuple, got.tolist())) == exp
    np.testing.assert_array_equal(orc.edge_overlap(A, B), got)
    # containment: B object strictly inside A object, no contact
    cont, _ = orc.predicate_join(A, B, 2)
    exp_c = set()
    for i in range(30):
        la, ha = box(cent[i], edges[i])
        for j in range(30):
            lb, hb = box(cent[30 + j], edges[30 + j])
            if np.all(lb > la) and np.all(hb < ha):
                exp_c.add((i, j))
    assert set(map(tuple, cont.tolist())) == exp_c


def test_directional_intersection_closed_form_and_invariants():
    """Intersection in one direction (pierce_intersection_main.cpp:103-155):
    edges of one side against the other's triangles, plus the first vertex of
    each object of that side strictly inside an object of the other.  On
    axis-aligned cubes the point test is a closed form (point strictly inside a
    box); the one-direction edge set is pinned by the union invariant of
    test_invariants_swap_self_and_predicates.  Both directions together give
    the undirected result, and swapping the sets swaps the directions."""
    A, B = nested_cubes()

    def first_vertices(m):
        out = {}
        for t, o in enumerate(m.triangle_to_object.tolist()):
            out.setdefault(o, m.vertices[m.indices[t, 0]])
        return out

    def boxes(m):
        res = {}
        for o in range(m.num_objects):
            v = m.vertices[np.unique(m.indices[m.triangle_to_object == o])]
            res[o] = (v.min(axis=0), v.max(axis=0))
        return res

    fa, fb, ba, bb = first_vertices(A), first_vertices(B), boxes(A), boxes(B)
    s = lambda x: set(map(tuple, x.tolist()))
    inside = lambda p, bx: bool(np.all(p > bx[0]) and np.all(p < bx[1]))
    d12 = s(orc.predicate_join(A, B, 1, direction=1)[0])
    d21 = s(orc.predicate_join(A, B, 1, direction=2)[0])
    exp12 = s(orc.edge_overlap(A, B, 1)) | {(i, j) for i in fa for j in bb if inside(fa[i], bb[j])}
    exp21 = s(orc.edge_overlap(A, B, 2)) | {(i, j) for j in fb for i in ba if inside(fb[j], ba[i])}
    assert d12 == exp12 and d21 == exp21
    assert d12 != d21 and len(d12) > 0 and len(d21) > 0
    assert d12 | d21 == s(orc.predicate_join(A, B, 1)[0])
    swapped = s(orc.predicate_join(B, A, 1, direction=2)[0])
    assert {(j, i) for i, j in swapped} == d12
    # containment ignores the direction
    np.testing.assert_array_equal(orc.predicate_join(A, B, 2, direction=1)[0], orc.predicate_join(A, B, 2)[0])


def test_bruteforce_matches_face_and_touch():
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=40, seed=11, level_min=1, level_max=1, box=12.0, jitter=0.02)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=40, seed=12, level_min=1, level_max=1, box=12.0, jitter=0.02)
    bf = orc.bruteforce_overlap(a, b)
    assert len(bf) > 20
    np.testing.assert_array_equal(orc.face_overlap(a, b), bf)
    np.testing.assert_array_equal(orc.edge_overlap(a, b), bf)
    assert orc.touch_overlap_count(a, b)[0] == len(bf)


def test_invariants_swap_self_and_predicates():
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=150, seed=21, box=25.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=150, seed=22, box=25.0, copy_fraction=0.3, copy_scale_min=0.3,
                    copy_scale_max=0.8, copy_source_count=150, copy_source_seed=21)
    ov = orc.face_overlap(a, b)
    sw = orc.face_overlap(b, a)
    np.testing.assert_array_equal(np.array(sorted(map(tuple, sw[:, ::-1].tolist()))).reshape(-1, 2), ov)
    self_pairs = set(map(tuple, orc.face_overlap(a, a).tolist()))
    assert all((i, i) in self_pairs for i in range(150))
    inter, _ = orc.predicate_join(a, b, 1)
    cont, _ = orc.predicate_join(a, b, 2)
    so, si, sc = set(map(tuple, ov.tolist())), set(map(tuple, inter.tolist())), set(map(tuple, cont.tolist()))
    assert so <= si and not (so & sc) and sc <= si
    assert len(sc) > 0
    d12 = set(map(tuple, orc.edge_overlap(a, b, 1).tolist()))
    d21 = set(map(tuple, orc.edge_overlap(a, b, 2).tolist()))
    assert d12 | d21 == so


def test_candidate_count_bruteforce():
    a = pz.generate(pz.GEN_ELLIPSOIDS, count=30, seed=31, level_min=1, level_max=1, box=10.0)
    b = pz.generate(pz.GEN_ELLIPSOIDS, count=30, seed=32, level_min=1, level_max=1, box=10.0)

    def count(S, T):
        lo = np.minimum(S.edge_starts, S.edge_ends)
        hi = np.maximum(S.edge_starts, S.edge_ends)
        tv = T.vertices[T.indices]
        tlo, thi = tv.min(1), tv.max(1)
        m = np.all(lo[:, None, :] <= thi[None], axis=2) & np.all(tlo[None] <= hi[:, None, :], axis=2)
        return int(m.sum())

    assert orc.count_candidate_tests(a, b) == count(a, b) + count(b, a)


def test_ray_target_hits_closed_form():
    # +Z ray from a point p against an axis-aligned cube [lo, hi] whose footprint
    # holds p strictly: one crossing per horizontal face above p, i.e.
    # (p.z < lo.z) + (p.z < hi.z); nothing when p's footprint is outside.  Target
    # faces sit on multiples of 0.5 and source corners never do, so no ray meets
    # a vertical face or a footprint boundary.
    rng = random.Random(21)
    tc = [np.array([rng.choice([0, 1, 2, 3]) for _ in range(3)], np.float32) for _ in range(25)]
    te = [rng.choice([1.0, 2.0, 4.0]) for _ in range(25)]
    sc = [np.array([rng.choice(np.arange(-1.0, 4.5, 0.5)) + 0.25 for _ in range(3)], np.float32) for _ in range(40)]
    T = cube_mesh(tc, te)
    S = cube_mesh(sc, [0.1] * len(sc))
    hits = orc.ray_target_hits(S, T)
    assert hits.shape == (40, 25)
    first = {}
    for t, o in enumerate(S.triangle_to_object.tolist()):
        first.setdefault(o, t)
    n_nonzero = 0
    for s in range(40):
        p = S.vertices[S.indices[first[s]][0]].astype(np.float64)
        for t in range(25):
            lo, hi = tc[t] - te[t] / 2, tc[t] + te[t] / 2
            inside_xy = lo[0] < p[0] < hi[0] and lo[1] < p[1] < hi[1]
            exp = int(inside_xy) * (int(p[2] < lo[2]) + int(p[2] < hi[2]))
            assert hits[s, t] == exp, (s, t)
            n_nonzero += exp > 0
            # odd count <=> the point is inside (the oracle's own containment test)
            assert (hits[s, t] % 2 == 1) == orc.point_in_object(T, t, p.astype(np.float32)), (s, t)
    assert n_nonzero > 20


def _grid(rows):
    g = np.zeros(len(rows), pz.GRID_DTYPE)
    for i, (ix, iy, iz, t, s, r) in enumerate(rows):
        g[i]["ix"], g[i]["iy"], g[i]["iz"] = ix, iy, iz
        g[i]["touch_count"], g[i]["avg_size_mean"], g[i]["vol_ratio"] = t, s, r
    return g


# Two grids whose whole-set estimate is exact in binary arithmetic.  a repeats
# cell (0,0,0): the first entry (touch 0) must lose to the second.  Means:
# size_a = (64*1 + 16*3 + 16*5) / 96 = 2, ratio_a = (64/16 + 16/4 + 16/4) / 96 =
# 1/8, so effective_a = 2 * cbrt(1/8) = 1; b is all size 1, ratio 1.
GRID_A = _grid([(0, 0, 0, 0, 100.0, 0.125), (0, 0, 0, 64, 1.0, 0.0625), (1, 0, 0, 16, 3.0, 0.25),
                (7, 7, 7, 16, 5.0, 0.25)])
GRID_B = _grid([(0, 0, 0, 8, 1.0, 1.0), (1, 0, 0, 16, 1.0, 1.0), (9, 9, 9, 5, 1.0, 1.0)])
# (cell size, epsilon, gamma, alpha on) -> (raw, alpha, final), by hand: cell
# (0,0,0) gives 64*8*min(1, (1+1+eps)^3/cs^3 * (sqrt(1/16))^gamma), cell (1,0,0)
# 16*16*min(1, (3+1+eps)^3/cs^3 * (sqrt(1/4))^gamma); alpha = max(1, (1+1)^3/cs^3).
GRID_CASES = [
    ((4.0, 0.0, 2.0, True), (4.0 + 64.0, 1.0, 68)),
    ((4.0, 0.0, 1.0, True), (16.0 + 128.0, 1.0, 144)),
    ((4.0, 2.0, 2.0, True), (32.0 + 216.0, 1.0, 248)),
    ((1.0, 0.0, 2.0, True), (256.0 + 256.0, 8.0, 64)),
    ((1.0, 0.0, 2.0, False), (256.0 + 256.0, 1.0, 512)),
]


@pytest.mark.parametrize("args,exp", GRID_CASES)
def test_grid_estimate_closed_form(args, exp):
    cs, eps, gamma, alpha = args
    got = orc.grid_estimate(GRID_A, GRID_B, cs, epsilon=eps, gamma=gamma, alpha_correction=alpha)
    assert got["has_grid"] and got["matched_cells"] == 2
    assert got["raw_pairs"] == exp[0] and got["alpha"] == exp[1] and got["estimated_pairs"] == exp[2]
    assert got["avg_size"] == (2.0, 1.0) and got["avg_vol_ratio"] == (0.125, 1.0)
    assert got["effective_size"] == (1.0, 1.0)
    assert not orc.grid_estimate(GRID_A, GRID_B[:0], cs)["has_grid"]
