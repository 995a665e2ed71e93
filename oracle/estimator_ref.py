"""CPU restatement of the learned output-sizing estimator.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker of the GPU
kernels in pierce_amd/csrc/kernels/estimator.hip.  Nothing shipped calls it.

The reference has no learned estimator.  It sizes buffers with the closed-form
grid estimate (pierce/query/src/cuda/estimated_overlap.cu:8-77, used from
pierce/query/src/applications/pierce_overlap_main.cpp:376-501).  This module
restates what the GPU side computes instead:

- ``object_boxes``: per-object boxes and triangle counts of a mesh (the
  statistics k_object_stats produces);
- ``features``: the 32 join features of k_feat_* (fp32 / fp64 operations in
  the kernels' order, integer fixed-point sums);
- ``forward`` / ``loss_grad``: the 32-128-128-2 ReLU network and the
  analytic gradient of the mean squared error, in float64;
- ``bf16`` / ``infer_bf16``: round-to-nearest-even bfloat16 and the
  bf16-operand, wide-accumulation inference of k_mlp_infer;
- ``adam_step``: one AdamW update.

Pinning (tests/test_estimator_ref.py):
- the gradient is checked against central finite differences;
- the forward pass reduces to a plain matrix product when every
  pre-activation is positive;
- bf16 rounding is checked on ties-to-even known answers;
- Adam's first step has the closed form lr * g / (|g| + eps);
- features are checked on hand-built box layouts whose values follow from
  the definitions.
"""
from __future__ import annotations

import numpy as np

F, H, O = 32, 128, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + O * H
P = OFF_B3 + O
GRID = 16
CELLS = GRID ** 3
BINS = 16
FIX = 65536.0


# ------------------------------------------------------------------ network
def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(O, H), p[OFF_B3:P])


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3])


def normalise(norm, X) -> np.ndarray:
    """(X - mean) * inv_std in float32, as the kernels compute it."""
    norm = np.asarray(norm, dtype=np.float32)
    X = np.asarray(X, dtype=np.float32).reshape(-1, F)
    return (X - norm[:F]) * norm[F:]


def forward(p, norm, X):
    """float64 forward; returns (Y, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xs = normalise(norm, X).astype(np.float64)
    z1 = xs @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ W3.T + b3
    return y, (xs, z1, a1, z2, a2)


def loss_grad(p, norm, X, T):
    """Loss = mean over samples of sum over outputs of (y - t)^2, and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (xs, z1, a1, z2, a2) = forward(p, norm, X)
    T = np.asarray(T, dtype=np.float64).reshape(-1, O)
    B = y.shape[0]
    e = y - T
    loss = float(np.mean(np.sum(e * e, axis=1)))
    dy = 2.0 * e / B
    gW3 = dy.T @ a2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3) * (z2 > 0)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ xs
    gb1 = dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, 0x7FC00000, r)
    return r.astype(np.uint32).view(np.float32)


def infer_bf16(p, norm, X) -> np.ndarray:
    """bf16 operands, exact products, wide accumulation, fp32 bias and ReLU,
    activations rounded to bf16 between layers."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    w1, w2, w3 = (bf16(w.astype(np.float32)).astype(np.float64) for w in (W1, W2, W3))
    xs = bf16(normalise(norm, X)).astype(np.float64)
    h1 = bf16(np.maximum((xs @ w1.T).astype(np.float32) + b1.astype(np.float32), 0)).astype(np.float64)
    h2 = bf16(np.maximum((h1 @ w2.T).astype(np.float32) + b2.astype(np.float32), 0)).astype(np.float64)
    return (h2 @ w3.T).astype(np.float32) + b3.astype(np.float32)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW update (decoupled weight decay); returns (p, m, v)."""
    p, g, m, v = (np.asarray(a, dtype=np.float64) for a in (p, g, m, v))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * (mh / (np.sqrt(vh) + eps) + wd * p), m, v


# ----------------------------------------------------------------- features
def object_boxes(mesh):
    """(boxes (n, 6) float32 lo xyz hi xyz, triangle counts (n,)) per object id."""
    V = np.asarray(mesh.vertices, dtype=np.float32).reshape(-1, 3)
    I = np.asarray(mesh.indices).reshape(-1, 3)
    t2o = np.asarray(mesh.triangle_to_object)
    n = int(t2o.max()) + 1 if t2o.size else 0
    tv = V[I]  # (nt, 3, 3)
    lo = np.full((n, 3), np.inf, np.float32)
    hi = np.full((n, 3), -np.inf, np.float32)
    np.minimum.at(lo, t2o, tv.min(axis=1))
    np.maximum.at(hi, t2o, tv.max(axis=1))
    cnt = np.bincount(t2o, minlength=n).astype(np.int64)
    return np.concatenate([lo, hi], axis=1), cnt


def _overlap_prob(d):
    return np.where(d >= 1.0, 1.0, d * (2.0 - d))


def features(boxA, triA, boxB, triB, predicates, partA=None, partB=None, nparts=1) -> np.ndarray:
    """32 features per part; predicates: one int per part (0, 1, 2)."""
    f32 = np.float32
    sets = []
    for box, tri, part in ((boxA, triA, partA), (boxB, triB, partB)):
        box = np.asarray(box, np.float32).reshape(-1, 6)
        tri = np.asarray(tri, np.int64)
        part = np.zeros(len(tri), np.int64) if part is None else np.asarray(part, np.int64)
        sets.append((box, tri, part))
    X = np.zeros((nparts, F), np.float32)
    preds = np.broadcast_to(np.asarray(predicates, np.int64), (nparts,))
    for p in range(nparts):
        acc = []
        for box, tri, part in sets:
            sel = (tri > 0) & (part == p)
            b = box[sel]
            acc.append(dict(box=b, tri=tri[sel], n=int(sel.sum()),
                            lo=b[:, :3].min(0) if len(b) else None, hi=b[:, 3:].max(0) if len(b) else None))
        A, B = acc
        f = np.zeros(F, np.float64)
        f[0] = np.log2(f32(1) + f32(A["n"]))
        f[1] = np.log2(f32(1) + f32(B["n"]))
        f[2] = np.log2(1.0 + float(A["tri"].sum()))
        f[3] = np.log2(1.0 + float(B["tri"].sum()))
        f[4] = np.log2(f32(1) + f32(A["tri"].max() if A["n"] else 0))
        f[5] = np.log2(f32(1) + f32(B["tri"].max() if B["n"] else 0))
        f[6 + int(preds[p])] = 1.0
        if A["n"] and B["n"]:
            ulo = np.minimum(A["lo"], B["lo"])
            uhi = np.maximum(A["hi"], B["hi"])
            rlo = np.maximum(A["lo"], B["lo"])
            rhi = np.minimum(A["hi"], B["hi"])
            r_empty = bool(np.any(rlo > rhi))
            cs = np.maximum((uhi - ulo) / f32(GRID), f32(1e-30)).astype(np.float32)
            cells, ext, n_in_r = [], [], []
            for S in (A, B):
                lo, hi = S["box"][:, :3], S["box"][:, 3:]
                c = (lo + hi) * f32(0.5)
                u = (c - ulo) / cs
                cell = np.clip(np.floor(u).astype(np.int64), 0, GRID - 1)
                e = np.minimum((hi - lo) / cs, f32(FIX))
                q = (e * f32(FIX) + f32(0.5)).astype(np.uint64)
                meets = np.all(~((lo > rhi) | (hi < rlo)), axis=1) & (not r_empty)
                flat = (cell[:, 2] * GRID + cell[:, 1]) * GRID + cell[:, 0]
                cnt = np.bincount(flat, minlength=CELLS).astype(np.float64)
                ce = np.zeros((CELLS, 3), np.uint64)
                np.add.at(ce, flat, q)
                cells.append((cnt, ce))
                ext.append(q.sum(axis=0, dtype=np.uint64))
                n_in_r.append(int(meets.sum()))
            f[9] = np.log2(1.0 + n_in_r[0])
            f[10] = np.log2(1.0 + n_in_r[1])
            # uniform estimate over the intersection box
            euni = 0.0
            if not r_empty:
                euni = float(n_in_r[0]) * float(n_in_r[1])
                for k in range(3):
                    ea = float(ext[0][k]) / A["n"] / FIX * float(cs[k])
                    eb = float(ext[1][k]) / B["n"] / FIX * float(cs[k])
                    side = float(rhi[k]) - float(rlo[k])
                    euni *= float(_overlap_prob(0.5 * (ea + eb) / side)) if side > 0 else 1.0
            f[11] = np.log2(1.0 + euni)
            (ca, ea_), (cb, eb_) = cells
            both = (ca > 0) & (cb > 0)
            hist = np.zeros(BINS + 2, np.uint64)
            for c in np.nonzero(both)[0]:
                est = ca[c] * cb[c]
                for k in range(3):
                    da = float(ea_[c, k]) / ca[c] / FIX
                    db = float(eb_[c, k]) / cb[c] / FIX
                    est *= float(_overlap_prob(0.5 * (da + db)))
                if est > 0:
                    b = min(max(int(np.floor(np.log2(est))) + 4, 0), BINS - 1)
                    qv = np.uint64(min(est, 2.0 ** 40) * FIX + 0.5)
                    hist[b] += qv
                    hist[BINS] += qv
                hist[BINS + 1] += np.uint64(1)
            f[12] = np.log2(1.0 + float(hist[BINS]) / FIX)
            for b in range(BINS):
                f[13 + b] = np.log2(1.0 + float(hist[b]) / FIX)
            f[29] = np.log2(1.0 + float(hist[BINS + 1]))
            ga = np.prod([float(ext[0][k]) / A["n"] / FIX for k in range(3)])
            gb = np.prod([float(ext[1][k]) / B["n"] / FIX for k in range(3)])
            f[30] = np.log2(np.cbrt(ga) + 2.0 ** -16)
            f[31] = np.log2(np.cbrt(gb) + 2.0 ** -16)
        X[p] = f.astype(np.float32)
    return X
