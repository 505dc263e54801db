"""The batch primitives on the GPU against numpy, for every built batch size:
the insert path's pad-and-sort (ibitonicSort, heap/heaputil.cuh:23-47) must
equal numpy's sort of the padded batch, and every merge kernel (imergePath,
heap/heaputil.cuh:99-158: register bitonic, merge path, group merge, serial
merge path, and the two halves taken one after the other as the heap's sift
steps take them) must split the sorted union at B."""
import numpy as np
import pytest

import gpuheap_amd as G
from gpuheap_amd import DeviceBuffer
from gpuheap_amd._lib import check, lib

pytestmark = pytest.mark.gpu

BATCHES = [64, 128, 256, 512, 1024, 2048, 4096, 8192]
KEYS = {"u32": (0, np.uint32), "u64": (1, np.uint64)}


def _rand(rng, dtype, shape, dup_every=3):
    info = np.iinfo(dtype)
    k = rng.integers(0, int(info.max), size=shape, dtype=np.uint64).astype(dtype)
    k[::dup_every] %= dtype(13)  # many duplicates
    return k


@pytest.mark.parametrize("B", BATCHES)
@pytest.mark.parametrize("key", ["u32", "u64"])
def test_sort_batch(gpu, B, key):
    kt, dtype = KEYS[key]
    if key == "u64" and B == 8192:
        pytest.skip("u64 keys stop at B=4096 (LDS)")
    rng = np.random.default_rng(B)
    nb = 64
    keys = _rand(rng, dtype, (nb, B))
    sizes = rng.integers(1, B + 1, size=nb).astype(np.int32)
    sizes[:4] = [B, 1, B - 1, B // 2]
    pad = int(np.iinfo(dtype).max)
    d_in, d_sz = DeviceBuffer.from_numpy(keys), DeviceBuffer.from_numpy(sizes)
    d_out = DeviceBuffer(keys.nbytes)
    check(lib().bgpq_test_sort(kt, B, d_in.ptr, d_sz.ptr, d_out.ptr, pad, nb))
    got = d_out.to_numpy(dtype).reshape(nb, B)
    for i in range(nb):
        want = np.full(B, pad, dtype=dtype)
        want[:sizes[i]] = keys[i, :sizes[i]]
        assert np.array_equal(got[i], np.sort(want)), f"batch {i} size {sizes[i]}"


@pytest.mark.parametrize("B", BATCHES)
@pytest.mark.parametrize("key", ["u32", "u64"])
@pytest.mark.parametrize("algo", [0, 1, 2, 3, 4, 5, 6, 7],
                         ids=["bitonic", "merge_path", "group", "serial", "halves", "lean", "halves_copy", "group_copy"])
def test_merge_batches(gpu, B, key, algo):
    kt, dtype = KEYS[key]
    if key == "u64" and B == 8192:
        pytest.skip("u64 keys stop at B=4096 (LDS)")
    rng = np.random.default_rng(B + 1)
    nb = 64
    a = np.sort(_rand(rng, dtype, (nb, B)), axis=1)
    b = np.sort(_rand(rng, dtype, (nb, B), dup_every=5), axis=1)
    # disjoint, interleaved and identical batches
    b[0] = a[0]
    b[1] = a[1] + dtype(1)
    a[2], b[2] = np.sort(np.arange(B, dtype=dtype)), np.sort(np.arange(B, 2 * B, dtype=dtype))
    a[3], b[3] = b[2].copy(), a[2].copy()
    top = np.iinfo(dtype).max  # keys equal to the padding value of the kernels
    a[4][-(B // 4):] = top
    b[4][-(B // 3):] = top
    # a node near the root against a batch of spread keys, a few of them
    # below the node's largest (mostly copy windows), both ways round, and
    # one key of the spread batch equal to the node's largest
    a[5] = np.arange(B, dtype=dtype) * dtype(3)
    few = np.array([1, 3 * B // 2, 3 * (B - 1)], dtype=dtype)
    b[5] = np.sort(np.concatenate([few, rng.integers(3 * B, int(top), B - 3, dtype=np.uint64).astype(dtype)]))
    a[6], b[6] = b[5].copy(), a[5].copy()
    a[7] = a[5]
    b[7] = np.sort(np.concatenate([a[5][-1:], rng.integers(3 * B, int(top), B - 1, dtype=np.uint64).astype(dtype)]))
    d = [DeviceBuffer.from_numpy(x) for x in (a, b)]
    small, large = DeviceBuffer(a.nbytes), DeviceBuffer(a.nbytes)
    check(lib().bgpq_test_merge(kt, B, d[0].ptr, d[1].ptr, small.ptr, large.ptr, nb, algo))
    s = small.to_numpy(dtype).reshape(nb, B)
    l_ = large.to_numpy(dtype).reshape(nb, B)
    for i in range(nb):
        u = np.sort(np.concatenate([a[i], b[i]]))
        assert np.array_equal(s[i], u[:B]), f"small half of batch {i}"
        assert np.array_equal(l_[i], u[B:]), f"large half of batch {i}"
