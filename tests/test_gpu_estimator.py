"""GPU estimator kernels against the CPU restatement (oracle/estimator_ref.py).

Checks:
- features from the object boxes (plain and batched datasets);
- bf16 matrix-core inference against bf16 emulation and the float64 network;
- fp32 forward, batch gradient and AdamW step against float64;
- graph-captured training against eager launches (bit-identical);
- convergence on a learnable target.
"""
import numpy as np
import pytest

import pierce_amd as pz
from oracle import estimator_ref as er
from pierce_amd.stream import translate

pytestmark = pytest.mark.gpu


def _rand_model(ctx, seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / er.F), (er.H, er.F))
    W2 = rng.normal(0, np.sqrt(2 / er.H), (er.H, er.H))
    W3 = rng.normal(0, np.sqrt(1 / er.H), (er.O, er.H))
    p = er.pack(W1, rng.normal(0, 0.1, er.H), W2, rng.normal(0, 0.1, er.H), W3, rng.normal(0, 0.1, er.O))
    p = p.astype(np.float32)
    norm = np.concatenate([rng.normal(0, 1, er.F), rng.uniform(0.5, 2, er.F)]).astype(np.float32)
    m = pz.Model(ctx, seed=1)
    m.set(params=p, norm=norm)
    return m, p, norm, rng


def _sets(seed, nA=150, nB=120, shift=(20.0, -10.0, 5.0)):
    kw = dict(box=80.0, tri_min=500, tri_median=1500, tri_max=4000)
    A = pz.generate(pz.GEN_QUERY_MIX, count=nA, seed=seed, **kw)
    B = translate(pz.generate(pz.GEN_QUERY_MIX, count=nB, seed=seed + 1, **kw), shift)
    return A, B


def test_join_features_match_reference(ctx):
    A, B = _sets(11)
    bA, tA = er.object_boxes(A)
    bB, tB = er.object_boxes(B)
    da, db = ctx.upload(A), ctx.upload(B)
    for pred in (0, 1, 2):
        print(f"[features] predicate {pred}", flush=True)
        got = pz.join_features(ctx, da, db, pred)
        want = er.features(bA, tA, bB, tB, [pred])[0]
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6)
    assert got[12] > 0 and got[29] > 0  # the sets share cells


def test_batch_features_match_reference(ctx):
    parts = [_sets(20 + 2 * q, nA=10 + 17 * q, nB=5 + 23 * q, shift=(3.0 * q, 0.0, -2.0)) for q in range(6)]
    parts.append((parts[0][0], translate(parts[0][1], (1e4, 0, 0))))  # disjoint query
    # objects without triangles inside a query (ids 2 and 5 of mesh1 empty)
    A, B = parts[3]
    keep = (A.triangle_to_object != 2) & (A.triangle_to_object != 5)
    parts.append((pz.extract_edges(pz.Mesh(A.vertices, A.indices[keep], A.triangle_to_object[keep])), B))
    preds = [q % 3 for q in range(len(parts))]
    da = ctx.upload_batch([p[0] for p in parts])
    db = ctx.upload_batch([p[1] for p in parts])
    got = pz.batch_features(ctx, da, db, preds)
    for q, (A, B) in enumerate(parts):
        bA, tA = er.object_boxes(A)
        bB, tB = er.object_boxes(B)
        want = er.features(bA, tA, bB, tB, [preds[q]])[0]
        np.testing.assert_allclose(got[q], want, rtol=2e-6, atol=2e-6, err_msg=f"query {q}")
    assert got[6][11] == 0 and got[6][12] == 0  # the disjoint query


@pytest.mark.parametrize("n", [1, 16, 37, 300])
def test_mfma_inference_matches_bf16_emulation(ctx, n):
    m, p, norm, rng = _rand_model(ctx, 5 + n)
    X = rng.normal(0, 2, (n, er.F)).astype(np.float32) * norm[er.F:] ** -1 + norm[:er.F]
    y = m.predict(X)
    yb = er.infer_bf16(p, norm, X)
    y64, _ = er.forward(p, norm, X)
    scale = 1 + np.abs(y64)
    # same bf16 operands: only accumulation order and rare one-ulp activation roundings differ
    assert np.max(np.abs(y - yb) / scale) < 2e-3, np.max(np.abs(y - yb) / scale)
    # against float64: within bf16 precision of the operands
    assert np.max(np.abs(y - y64) / scale) < 5e-2
    m.close()


def test_fp32_forward_matches_float64(ctx):
    m, p, norm, rng = _rand_model(ctx, 7)
    X = rng.normal(0, 1, (50, er.F)).astype(np.float32) / norm[er.F:] + norm[:er.F]
    y = m.forward(X)
    y64, _ = er.forward(p, norm, X)
    np.testing.assert_allclose(y, y64, rtol=1e-4, atol=1e-4)
    m.close()


def test_gradients_match_reference(ctx):
    m, p, norm, rng = _rand_model(ctx, 8)
    X = rng.normal(0, 1, (64, er.F)).astype(np.float32) / norm[er.F:] + norm[:er.F]
    T = rng.normal(0, 3, (64, er.O)).astype(np.float32)
    loss, g = m.gradients(X, T)
    loss64, g64 = er.loss_grad(p, norm, X, T)
    assert loss == pytest.approx(loss64, rel=1e-5)
    # fp32 accumulation over 64 samples: errors relative to the gradient scale of each block
    for lo, hi in ((er.OFF_W1, er.OFF_B1), (er.OFF_B1, er.OFF_W2), (er.OFF_W2, er.OFF_B2), (er.OFF_B2, er.OFF_W3),
                   (er.OFF_W3, er.OFF_B3), (er.OFF_B3, er.P)):
        err = np.max(np.abs(g[lo:hi] - g64[lo:hi]))
        assert err <= 1e-4 * np.max(np.abs(g64[lo:hi])) + 1e-7, (lo, err)
    m.close()


def test_adam_step_matches_reference(ctx):
    m, p, norm, rng = _rand_model(ctx, 9)
    n = 48
    X = rng.normal(0, 1, (n, er.F)).astype(np.float32) / norm[er.F:] + norm[:er.F]
    T = rng.normal(0, 3, (n, er.O)).astype(np.float32)
    _, g64 = er.loss_grad(p, norm, X, T)
    lr = 1e-3
    m.train(X, T, epochs=1, batch=n, lr=lr, lr_final=lr, use_graph=0, validation_fraction=0.0)
    p1, _, _ = m.get()
    z = np.zeros(er.P)
    want, _, _ = er.adam_step(p.astype(np.float64), g64, z, z, 1, lr=lr)
    sure = np.abs(g64) > 1e-3 * np.max(np.abs(g64))  # step 1 moves by lr * sign(g): skip near-zero gradients
    assert sure.sum() > 0.5 * er.P
    np.testing.assert_allclose(p1[sure], want[sure], rtol=0, atol=4e-7)
    m.close()


def test_graph_training_matches_eager(ctx):
    rng = np.random.default_rng(10)
    X = rng.normal(0, 1, (100, er.F)).astype(np.float32)
    T = np.stack([X[:, :4].sum(1), np.abs(X[:, 4:8]).sum(1)], axis=1).astype(np.float32)
    res = []
    for use_graph in (1, 0):
        m = pz.Model(ctx, seed=3)
        losses = m.train(X, T, epochs=4, batch=16, use_graph=use_graph, seed=5)
        res.append((m.get()[0], losses, m.get()[2]))
        m.close()
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    np.testing.assert_array_equal(res[0][2], res[1][2])


def test_training_converges_and_saves(ctx, tmp_path):
    rng = np.random.default_rng(11)
    X = rng.normal(0, 1, (512, er.F)).astype(np.float32)
    T = np.stack([2 * X[:, 0] - X[:, 1] + 0.5 * X[:, 2] * X[:, 3], np.maximum(X[:, 4], 0) * 3],
                 axis=1).astype(np.float32)
    m = pz.Model(ctx, seed=2)
    losses = m.train(X, T, epochs=150, batch=64)
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
    y = m.predict(X)
    assert np.mean((y - T) ** 2) < 0.1
    path = str(tmp_path / "model.pzm")
    m.save(path)
    m2 = pz.Model(ctx, path=path)
    for a, b in zip(m.get(), m2.get()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(m2.predict(X), y)
    m.close()
    m2.close()
