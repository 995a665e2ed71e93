"""Pinning tests of the estimator's CPU restatement (oracle/estimator_ref.py).

Each check is chosen so that a plausible mistake fails it:
- finite differences catch a dropped term, a wrong sign or a transposed
  operand in the backward pass;
- the linear reduction catches forward-pass mistakes;
- known answers pin the bf16 rounding and Adam's first step;
- hand-built box layouts pin the features.
"""
import numpy as np
import pytest

from oracle import estimator_ref as er


def _rand_params(rng, scale=1.0):
    W1 = rng.normal(0, scale * np.sqrt(2 / er.F), (er.H, er.F))
    W2 = rng.normal(0, scale * np.sqrt(2 / er.H), (er.H, er.H))
    W3 = rng.normal(0, scale * np.sqrt(1 / er.H), (er.O, er.H))
    return er.pack(W1, rng.normal(0, 0.1, er.H), W2, rng.normal(0, 0.1, er.H), W3, rng.normal(0, 0.1, er.O))


def _norm(rng):
    return np.concatenate([rng.normal(0, 1, er.F), rng.uniform(0.5, 2, er.F)]).astype(np.float32)


def test_layout_offsets():
    assert (er.OFF_B1, er.OFF_W2, er.OFF_B2, er.OFF_W3, er.OFF_B3, er.P) == (4096, 4224, 20608, 20736, 20992, 20994)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _rand_params(rng)
    norm = _norm(rng)
    X = rng.normal(0, 2, (9, er.F)).astype(np.float32)
    T = rng.normal(0, 3, (9, er.O))
    _, g = er.loss_grad(p, norm, X, T)
    h = 1e-6
    # parameters from every block, including both output units
    idx = np.concatenate([rng.choice(np.arange(lo, hi), size=12, replace=False)
                          for lo, hi in ((er.OFF_W1, er.OFF_B1), (er.OFF_B1, er.OFF_W2), (er.OFF_W2, er.OFF_B2),
                                         (er.OFF_B2, er.OFF_W3), (er.OFF_W3, er.OFF_B3))] + [np.arange(er.OFF_B3, er.P)])
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (er.loss_grad(pp, norm, X, T)[0] - er.loss_grad(pm, norm, X, T)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-5 * max(1.0, abs(num)), (i, num, g[i])


def test_forward_reduces_to_linear_map_when_all_units_active():
    rng = np.random.default_rng(4)
    W1 = rng.uniform(0, 1, (er.H, er.F))
    W2 = rng.uniform(0, 1, (er.H, er.H))
    W3 = rng.normal(0, 1, (er.O, er.H))
    b1, b2, b3 = rng.uniform(0, 1, er.H), rng.uniform(0, 1, er.H), rng.normal(0, 1, er.O)
    p = er.pack(W1, b1, W2, b2, W3, b3)
    norm = np.concatenate([np.zeros(er.F), np.ones(er.F)]).astype(np.float32)
    X = rng.uniform(0, 1, (5, er.F)).astype(np.float32)
    y, _ = er.forward(p, norm, X)
    A = W3 @ W2 @ W1
    c = W3 @ (W2 @ b1 + b2) + b3
    np.testing.assert_allclose(y, X.astype(np.float64) @ A.T + c, rtol=1e-12)


def test_forward_zero_weights_gives_bias():
    p = np.zeros(er.P)
    p[er.OFF_B3:] = [3.5, -1.25]
    norm = np.concatenate([np.zeros(er.F), np.ones(er.F)]).astype(np.float32)
    y, _ = er.forward(p, norm, np.ones((3, er.F), np.float32))
    np.testing.assert_array_equal(y, [[3.5, -1.25]] * 3)


def test_bf16_rounding_known_answers():
    ulp = 2.0 ** -7
    x = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + ulp / 2 + 2 ** -20, -2.0, 0.0, 3.0e38], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -2.0, 0.0, np.float32(3.0e38)], np.float64)
    got = er.bf16(x).astype(np.float64)
    np.testing.assert_array_equal(got[:6], want[:6])
    assert got[6] > 2.9e38 and np.isfinite(got[6])
    assert np.isnan(er.bf16(np.array([np.nan], np.float32))[0])


def test_infer_bf16_close_to_float64_forward():
    rng = np.random.default_rng(5)
    p = _rand_params(rng)
    norm = _norm(rng)
    X = rng.normal(0, 1, (20, er.F)).astype(np.float32)
    y64, _ = er.forward(p, norm, X)
    yb = er.infer_bf16(p, norm, X)
    assert np.max(np.abs(yb - y64)) < 0.05 * (1 + np.max(np.abs(y64)))


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    p = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    z = np.zeros(50)
    p1, m1, v1 = er.adam_step(p, g, z, z, 1, lr=1e-3)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)
    # decoupled weight decay shrinks every parameter by lr * wd * p on top
    p2, _, _ = er.adam_step(p, g, z, z, 1, lr=1e-3, wd=0.5)
    np.testing.assert_allclose(p2, p1 - 1e-3 * 0.5 * p, rtol=1e-12)


def _box(lo, hi):
    return np.array(list(lo) + list(hi), np.float32)


def test_features_two_boxes_closed_form():
    # A = [0,1]^3, B = [0.5,1.5]^3: union side 1.5, cell 1.5/16, centres in cells 5 and 10.
    A = _box((0, 0, 0), (1, 1, 1))[None]
    B = _box((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))[None]
    X = er.features(A, [12], B, [80], [2])[0]
    cs = 1.5 / 16
    assert X[0] == 1.0 and X[1] == 1.0
    assert X[2] == pytest.approx(np.log2(13)) and X[3] == pytest.approx(np.log2(81))
    assert X[4] == pytest.approx(np.log2(13)) and X[5] == pytest.approx(np.log2(81))
    assert list(X[6:9]) == [0, 0, 1]
    assert X[9] == 1.0 and X[10] == 1.0                  # both boxes meet the intersection
    assert X[11] == pytest.approx(1.0)                   # extents cover the intersection: one expected pair
    assert X[12] == 0 and not X[13:29].any() and X[29] == 0  # no cell holds both sets
    assert X[30] == pytest.approx(np.log2(1 / cs + 2 ** -16), rel=1e-6)
    assert X[31] == pytest.approx(X[30], rel=1e-6)


def test_features_same_cell_pair_probability():
    # Two tiny boxes in one cell: per-axis overlap probability d(2 - d) with
    # d = mean extent in cells; the pair lands in bin floor(log2(est)) + 4.
    scene = _box((0, 0, 0), (16, 16, 16))  # cell size 1 once both corners are in the union
    A = np.stack([scene, _box((3.2, 3.2, 3.2), (3.4, 3.4, 3.4))])
    B = np.stack([_box((3.5, 3.5, 3.5), (3.7, 3.7, 3.7)), scene])
    X = er.features(A, [12, 12], B, [12, 12], [0])[0]
    # cell (3,3,3) holds one small box of each set; cell (8,8,8) holds both scene boxes
    d = 0.2  # extent in cells of each small box (float32 rounding below 1e-6)
    est_small = (d * (2 - d)) ** 3
    est_scene = 1.0
    total = est_small + est_scene
    assert X[12] == pytest.approx(np.log2(1 + total), rel=1e-5)
    small_bin = max(int(np.floor(np.log2(est_small))) + 4, 0)  # log2(0.047) = -4.4 -> clamped to bin 0
    assert small_bin == 0
    assert X[13 + small_bin] == pytest.approx(np.log2(1 + est_small), rel=5e-4)  # 2^-16 fixed-point extents
    assert X[13 + 4] == pytest.approx(1.0)  # est 1 -> bin 4 holds log2(1 + 1)
    assert X[29] == pytest.approx(np.log2(3))


def test_features_disjoint_sets_have_no_overlap_terms():
    A = _box((0, 0, 0), (1, 1, 1))[None]
    B = _box((5, 5, 5), (6, 6, 6))[None]
    X = er.features(A, [12], B, [12], [1])[0]
    assert X[9] == 0 and X[10] == 0 and X[11] == 0 and X[12] == 0
    assert X[7] == 1


def test_features_one_empty_side():
    A = _box((0, 0, 0), (1, 1, 1))[None]
    X = er.features(A, [12], np.zeros((0, 6), np.float32), np.zeros(0, np.int64), [0])[0]
    assert X[0] == 1 and X[1] == 0 and not X[9:].any()
