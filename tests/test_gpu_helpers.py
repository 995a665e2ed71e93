"""The reference's extern "C" helpers through the C ABI on device buffers:
merge_and_deduplicate_pairs_gpu and compact_hash_table_pairs
(pierce/query/src/cuda/mesh_query_deduplication.h:6-18), exclusive_scan_gpu
(pierce/query/src/cuda/scan_utils.h:7-9) and the GridCell-layout
estimateOverlapSelectivity (pierce/query/src/cuda/estimated_overlap.h:6-13).
Device memory comes from torch; the checks are numpy sort/unique/cumsum and a
float32 evaluation of the per-cell formula."""
import ctypes as C

import numpy as np
import pytest

import pierce_amd as pz

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a: np.ndarray):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def pairs_of(t, n):
    torch.cuda.synchronize()
    return t.cpu().numpy().reshape(-1, 2)[:n]


def unique_pairs(p):
    p = np.asarray(p, np.int32).reshape(-1, 2)
    if len(p) == 0:
        return p
    return np.unique(p, axis=0)  # lexicographic (mesh1, mesh2)


def random_pairs(rng, n, hi=50):
    return rng.integers(0, hi, size=(n, 2)).astype(np.int32)


@pytest.fixture(scope="module")
def lib(ctx):  # the context makes device 0 current
    return pz.lib()


@pytest.mark.parametrize("n1,n2", [(0, 0), (1, 0), (0, 7), (1000, 513), (70000, 40000)])
def test_merge_two_arrays(lib, n1, n2):
    rng = np.random.default_rng(n1 + 3 * n2)
    p1, p2 = random_pairs(rng, n1), random_pairs(rng, n2, hi=300)
    d1, d2 = dev(p1 if n1 else np.zeros((1, 2), np.int32)), dev(p2 if n2 else np.zeros((1, 2), np.int32))
    out = dev(np.full((max(n1 + n2, 1), 2), -7, np.int32))
    print(f"[merge] n1={n1} n2={n2}", flush=True)
    n = lib.merge_and_deduplicate_pairs_gpu(d1.data_ptr() if n1 else None, n1, d2.data_ptr() if n2 else None, n2,
                                            out.data_ptr())
    exp = unique_pairs(np.concatenate([p1, p2]))
    assert n == len(exp)
    np.testing.assert_array_equal(pairs_of(out, n), exp)


def test_merge_in_place_forms(lib):
    """pierce_overlap_two_pass_main.cpp:137 passes (nullptr, total, nullptr, 0, out):
    the pairs are already in the output buffer."""
    rng = np.random.default_rng(11)
    allp = random_pairs(rng, 5000, hi=40)
    out = dev(allp)
    n = lib.merge_and_deduplicate_pairs_gpu(None, len(allp), None, 0, out.data_ptr())
    exp = unique_pairs(allp)
    assert n == len(exp)
    np.testing.assert_array_equal(pairs_of(out, n), exp)
    # second part in place behind a separately passed first part
    p1, p2 = random_pairs(rng, 300, hi=30), random_pairs(rng, 900, hi=30)
    out = dev(np.concatenate([np.zeros_like(p1), p2]))
    d1 = dev(p1)
    n = lib.merge_and_deduplicate_pairs_gpu(d1.data_ptr(), len(p1), None, len(p2), out.data_ptr())
    exp = unique_pairs(np.concatenate([p1, p2]))
    assert n == len(exp)
    np.testing.assert_array_equal(pairs_of(out, n), exp)
    # invalid counts are rejected, no work done
    assert lib.merge_and_deduplicate_pairs_gpu(None, -1, None, 0, out.data_ptr()) == -1


@pytest.mark.parametrize("cap", [1 << 20, 37])
def test_compact_hash_table(lib, cap):
    rng = np.random.default_rng(5)
    size = 1 << 16
    table = np.full(size, np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64)
    keys = np.unique(rng.integers(0, 1 << 40, size=3000, dtype=np.uint64))
    slots = rng.choice(size, size=len(keys), replace=False)
    table[slots] = keys
    dt = dev(table.view(np.int64))
    out = dev(np.zeros((len(keys), 2), np.int32))
    n = lib.compact_hash_table_pairs(dt.data_ptr(), size, out.data_ptr(), cap)
    assert n == min(len(keys), cap)
    got = pairs_of(out, n).astype(np.int64)
    got_keys = (got[:, 0].astype(np.uint64) << np.uint64(32)) | got[:, 1].astype(np.uint32).astype(np.uint64)
    assert len(set(got_keys.tolist())) == n
    assert set(got_keys.tolist()) <= set(keys.tolist())
    if cap >= len(keys):
        assert set(got_keys.tolist()) == set(keys.tolist())


@pytest.mark.parametrize("n", [0, 1, 1000, 3_000_001])
def test_exclusive_scan(lib, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, size=max(n, 1)).astype(np.int32)
    dx = dev(x)
    out = dev(np.full(max(n, 1), -1, np.int64))
    total = lib.exclusive_scan_gpu(dx.data_ptr() if n else None, out.data_ptr() if n else None, n)
    exp = np.concatenate([[0], np.cumsum(x[:n].astype(np.int64))])
    assert total == exp[n]
    if n:
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy()[:n], exp[:n])


def cell_formula(a, b, cell_volume, eps=np.float32(0.001), gamma=np.float32(0.8)):
    """estimated_overlap.cu:29-40 per matched cell in float32, summed in float64."""
    f = np.float32
    tot = 0.0
    for x, y in zip(a, b):
        if x["touch_count"] == 0 or y["touch_count"] == 0:
            continue
        comb = f(f(x["avg_size_mean"] + y["avg_size_mean"]) + eps)
        prob = f(f(f(comb * comb) * comb) / f(cell_volume))
        prob = f(prob * f(np.power(f(np.sqrt(f(x["vol_ratio"] * y["vol_ratio"]))), gamma)))
        prob = min(prob, f(1.0))
        tot += float(f(f(f(x["touch_count"]) * f(y["touch_count"])) * prob))
    return tot


@pytest.mark.parametrize("n", [0, 1, 5000])
def test_selectivity_on_gridcell_layout(lib, n):
    rng = np.random.default_rng(2 + n)
    a = np.zeros(n, pz.CELL_DTYPE)
    b = np.zeros(n, pz.CELL_DTYPE)
    for c in (a, b):
        c["center_count"] = rng.integers(0, 5, n)
        c["touch_count"] = rng.integers(0, 20, n)
        c["avg_size_mean"] = rng.uniform(0.1, 4.0, n).astype(np.float32)
        c["vol_ratio"] = rng.uniform(0.05, 1.0, n).astype(np.float32)
    got = pz.estimate_overlap_selectivity_cells(a, b, 125.0)
    exp = cell_formula(a, b, 125.0)
    assert got == pytest.approx(exp, rel=2e-6, abs=1e-6)
    # the 28-byte .pre cells through the host entry point give the same sum
    if n:
        g = np.zeros(n, pz.GRID_DTYPE)
        h = np.zeros(n, pz.GRID_DTYPE)
        for src, dst in ((a, g), (b, h)):
            for k in pz.CELL_DTYPE.names:
                dst[k] = src[k]
        assert pz.estimate_overlap_selectivity(g, h, 125.0) == pytest.approx(got, rel=2e-6)
    assert lib.pz_estimate_overlap_selectivity_cells(None, None, -1, C.c_float(1.0), C.c_float(0.001),
                                                     C.c_float(0.8)) == -1.0
