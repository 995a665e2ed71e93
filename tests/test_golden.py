"""Golden fixtures (tests/golden): brute-force pair sets of small seeded sets.

On the CPU: the oracle's BVH paths reproduce the brute-force sets, the
generators reproduce the stored .pre bytes, and the .pre reader/writer round
trip is byte-exact.
"""
import os

import numpy as np
import pytest

import pierce_amd as pz
from oracle import oracle as orc
from tests.golden.make_golden import cases
from tests.golden_io import CASES, GOLDEN, PREDICATES, expected, meshes


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_golden(case):
    a, b = meshes(case)
    np.testing.assert_array_equal(orc.face_overlap(a, b), expected(case, "overlap"))
    for name in ("intersection", "containment"):
        np.testing.assert_array_equal(orc.predicate_join(a, b, PREDICATES[name])[0], expected(case, name))


def test_golden_sets_are_nonempty_and_consistent():
    for case in CASES:
        ov = set(map(tuple, expected(case, "overlap").tolist()))
        it = set(map(tuple, expected(case, "intersection").tolist()))
        ct = set(map(tuple, expected(case, "containment").tolist()))
        assert ov, case
        assert ov <= it and ct <= it and not (ov & ct)  # intersection covers both; containment excludes contact
    assert len(expected("ellipsoids", "containment")) > 0


@pytest.mark.parametrize("case", CASES)
def test_generators_reproduce_fixture_bytes(case, tmp_path):
    m1, m2 = cases()[case]
    for i, m in ((1, m1), (2, m2)):
        p = str(tmp_path / f"{case}_{i}.pre")
        pz.write_pre(p, m)
        with open(p, "rb") as f, open(os.path.join(GOLDEN, f"{case}_mesh{i}.pre"), "rb") as g:
            assert f.read() == g.read()


def test_pre_round_trip_is_byte_exact(tmp_path):
    src = os.path.join(GOLDEN, "mix_mesh1.pre")
    p = str(tmp_path / "rt.pre")
    pz.write_pre(p, pz.read_pre(src))
    with open(p, "rb") as f, open(src, "rb") as g:
        assert f.read() == g.read()
