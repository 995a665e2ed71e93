"""merge_and_deduplicate_pairs_gpu_noalloc (fixed device scratch, no
allocation) against merge_and_deduplicate_pairs_gpu (growable workspace)
through the C ABI, on two calls in a row, the second with a smaller input, and
in the reference's in-place forms (pierce_overlap_two_pass_main.cpp:137)."""
import numpy as np
import pytest

import pierce_amd as pz

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def host_pairs(t, n):
    torch.cuda.synchronize()
    return t.cpu().numpy().reshape(-1, 2)[:n]


@pytest.fixture(scope="module")
def lib():
    ctx = pz.Context(0)  # makes device 0 current
    yield pz.lib()
    ctx.close()


def both(lib, p1, p2, in_place):
    """(result of the workspace entry point, result of the no-alloc one)."""
    out = []
    for fn in (lib.merge_and_deduplicate_pairs_gpu, lib.merge_and_deduplicate_pairs_gpu_noalloc):
        if in_place:  # everything already in the output buffer
            o = dev(np.concatenate([p1, p2]))
            n = fn(None, len(p1), None, len(p2), o.data_ptr())
        else:
            d1, d2 = dev(p1), dev(p2)
            o = dev(np.full((len(p1) + len(p2), 2), -3, np.int32))
            n = fn(d1.data_ptr(), len(p1), d2.data_ptr(), len(p2), o.data_ptr())
        assert n >= 0, pz.lib().pz_last_error()
        out.append(host_pairs(o, n))
    return out


@pytest.mark.parametrize("in_place", [False, True])
def test_no_alloc_matches_workspace_entry_point(lib, in_place):
    rng = np.random.default_rng(7 + in_place)
    # first call: near the fixed limit; second: smaller, so stale scratch would show
    sizes = [(600_000, pz.NOALLOC_MAX_PAIRS - 600_000), (3_000, 1_500)]
    for k, (n1, n2) in enumerate(sizes):
        p1 = rng.integers(0, 2000, size=(n1, 2)).astype(np.int32)
        p2 = rng.integers(0, 2000, size=(n2, 2)).astype(np.int32)
        print(f"[no-alloc] call {k}: n1={n1} n2={n2} in_place={in_place}", flush=True)
        ws, na = both(lib, p1, p2, in_place)
        exp = np.unique(np.concatenate([p1, p2]), axis=0)
        np.testing.assert_array_equal(ws, exp)
        np.testing.assert_array_equal(na, ws)


def test_no_alloc_limits(lib):
    big = dev(np.zeros((2, 2), np.int32))
    # one pair over the limit: rejected before any device work, with a message
    assert lib.merge_and_deduplicate_pairs_gpu_noalloc(None, pz.NOALLOC_MAX_PAIRS, None, 1, big.data_ptr()) == -1
    assert b"limit" in lib.pz_last_error()
    assert lib.merge_and_deduplicate_pairs_gpu_noalloc(None, -1, None, 0, big.data_ptr()) == -1
    assert lib.merge_and_deduplicate_pairs_gpu_noalloc(None, 0, None, 0, big.data_ptr()) == 0
    # and a valid call still works afterwards
    p = np.array([[3, 4], [1, 2], [3, 4]], np.int32)
    d = dev(p)
    o = dev(np.zeros((3, 2), np.int32))
    assert lib.merge_and_deduplicate_pairs_gpu_noalloc(d.data_ptr(), 3, None, 0, o.data_ptr()) == 2
    np.testing.assert_array_equal(host_pairs(o, 2), [[1, 2], [3, 4]])
