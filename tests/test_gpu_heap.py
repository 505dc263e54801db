"""GPU parity of the heap against the CPU oracle, through the C ABI.

* One workgroup (grid=1) performs the oracle model's data movement exactly:
  the whole device state (every slot, every counter) must equal the model's
  after every operation, for all four variants and both key types.
* Many workgroups: the reference's own harness (test/heap_test.cu,
  test/partial_heap_test.cu) -- insert half, concurrent insert+delete, delete
  the rest -- must emit numpy's sort of the keys, bit for bit, with
  checkInsertHeap holding after the insertion phase.
"""
import numpy as np
import pytest

import _oracle as O
import gpuheap_amd as G
from gpuheap_amd import DeviceBuffer, Heap, batch_num_for

pytestmark = pytest.mark.gpu

VARIANTS = list(G.VARIANTS)
PB = ("buitdd", "tditdd")
INIT = {"u32": O.U32_MAX, "u64": O.U64_MAX}


def _state_equal(h, m, key):
    items, status, ctr = h.snapshot()
    bc, pbs, dc, err = m.counters()
    assert ctr[3] == 0 and err == 0, f"error flags gpu={ctr[3]} model={err}"
    assert (ctr[0], ctr[1], ctr[2]) == (bc, pbs, dc)
    assert not status.any(), "a status word is not AVAIL after the operation"
    want = m.items()
    got = items.astype(np.uint64)
    if not np.array_equal(got, want):
        bad = np.nonzero(got != want)[0]
        raise AssertionError(f"slot mismatch at keys {bad[:8]} (slot {bad[0] // m.B})")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [64, 128, 256, 512, 1024, 2048, 4096, 8192])
@pytest.mark.parametrize("key", ["u32", "u64"])
def test_single_workgroup_state_parity(gpu, variant, B, key):
    if key == "u64" and B == 8192:
        pytest.skip("u64 keys stop at B=4096 (LDS); the rejection is tested below")
    print(f"[case] state parity {variant} B={B} {key}", flush=True)
    rng = np.random.default_rng(B * 7 + VARIANTS.index(variant))
    batch_num = 128
    h = Heap(batch_num, B, variant=variant, key=key)
    m = O.Model(variant, batch_num, B, INIT[key])
    dtype = np.uint32 if key == "u32" else np.uint64
    cap = 4096 * B
    d_out = DeviceBuffer(cap * np.dtype(dtype).itemsize)
    held = 0
    for step in range(80):
        if held == 0 or (held < 40 * B and rng.random() < 0.6):
            size = int(rng.integers(1, B + 1)) if variant in PB else B
            hi = 2**32 - 1 if key == "u32" else 2**64 - 1
            keys = rng.integers(0, hi, size=size, dtype=np.uint64)
            if step % 5 == 0:
                keys %= np.uint64(9)  # duplicates
            keys = keys.astype(dtype)
            d = DeviceBuffer.from_numpy(keys)
            h.insert(d, size, chunk=size, grid=1)
            assert m.insert(keys.astype(np.uint64)) == 0
            held += size
        else:
            h.delete(d_out, 1, grid=1)
            held -= m.delete()
        _state_equal(h, m, key)
    dc = m.counters()[2]
    got = d_out.to_numpy(dtype, dc).astype(np.uint64)
    assert np.array_equal(got, m.out[:dc])


def _harness(variant, B, n, key, seed, grid, distinct=0):
    """test/heap_test.cu main (:85-172) and test/partial_heap_test.cu: returns
    (heap output, expected, check-after-insert).  distinct > 0 folds each
    half of the u32 keys onto that many values (ties everywhere), keeping the
    harness's split: every key of the second half, which is inserted during
    the concurrent phase, stays above every key of the first."""
    if key == "u32":
        keys = O.harness_keys(seed, n)
        if distinct:
            keys = (keys & np.uint32(1 << 31)) | (keys % np.uint32(distinct))
    else:
        k32 = O.harness_keys(seed, n).astype(np.uint64)
        keys = (k32 << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    half = n // 2
    bn = batch_num_for(n, B)
    h = Heap(bn, B, variant=variant, key=key)
    pb = variant in PB
    chunk1 = max(1, B // 3) if pb else B
    chunk2 = max(1, 2 * B // 3) if pb else B
    d_in = DeviceBuffer.from_numpy(keys)
    d_out = DeviceBuffer(keys.nbytes)
    h.insert(d_in, half, chunk=chunk1, grid=grid)
    ok = h.checkInsertHeap()
    msg = getattr(h, "last_check_message", "")
    calls = (half + B - 1) // B
    second = DeviceBuffer.from_numpy(keys[half:])
    h.concurrent(second, n - half, d_out, calls, chunk=chunk2, grid=grid)
    h.delete(d_out, calls + 1, grid=grid)
    G.sync()
    assert h.error_flags() == 0, f"device error flags {h.error_flags()}"
    assert h.isEmpty()
    return d_out.to_numpy(keys.dtype, n), np.sort(keys), (ok, msg)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [128, 256, 512, 1024])
def test_reference_harness_matrix(gpu, variant, B):
    # test/test.sh:5-34: batch 128..1024, 2^20 keys (2^21, 2^22 in the slow test)
    print(f"[case] harness {variant} B={B}", flush=True)
    got, want, (ok, msg) = _harness(variant, B, 1 << 20, "u32", seed=B, grid=32)
    assert ok, msg
    assert np.array_equal(got, want)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [2048, 4096, 8192])
def test_large_batch_harness(gpu, variant, B):
    """The same harness with the batch sizes that need more than the 64 KB of
    LDS a workgroup gets by default (B=8192: 96 KB)."""
    print(f"[case] harness large batch {variant} B={B}", flush=True)
    got, want, (ok, msg) = _harness(variant, B, 1 << 21, "u32", seed=B + 1, grid=32)
    assert ok, msg
    assert np.array_equal(got, want)


def test_u64_batch_8192_rejected():
    # 3 * 8192 * 8 bytes of delete workspace do not fit in 160 KB of LDS
    with pytest.raises(G.BgpqError, match="LDS"):
        Heap(4, 8192, key="u64")


@pytest.mark.parametrize("variant", VARIANTS)
def test_reference_harness_u64(gpu, variant):
    print(f"[case] harness u64 {variant}", flush=True)
    got, want, (ok, msg) = _harness(variant, 512, 1 << 19, "u64", seed=3, grid=64)
    assert ok, msg
    assert np.array_equal(got, want)


def test_config1_buitdd_heapsort(gpu):
    """2^20 uniform u32 keys (SplitMix64 seed 0), buitdd, B=1024,
    batch_num 2048: insert all in B-key calls, checkInsertHeap, delete until
    empty; output equals the sorted keys."""
    print("[case] config1", flush=True)
    n, B = 1 << 20, 1024
    keys = O.keys_u32(0, n)
    h = Heap(batch_num_for(n, B), B, variant="buitdd")
    assert h.batch_num == 2048
    d = DeviceBuffer.from_numpy(keys)
    h.insert(d, n, chunk=B)
    assert h.checkInsertHeap(), h.last_check_message
    assert h.itemCount() == n and h.nodeCount() == n // B
    out = DeviceBuffer(keys.nbytes)
    h.delete(out, n // B + 1)
    G.sync()
    assert h.isEmpty() and h.error_flags() == 0
    assert np.array_equal(out.to_numpy(np.uint32, n), np.sort(keys))


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("order", ["random", "descending"])
def test_heapsort_api(gpu, variant, order):
    print(f"[case] heapsort {variant} {order}", flush=True)
    n, B = 1 << 18, 512
    keys = O.keys_u32(9, n)
    if order == "descending":
        keys = np.sort(keys)[::-1].copy()
    h = Heap(batch_num_for(n, B), B, variant=variant)
    d = DeviceBuffer.from_numpy(keys)
    t_ins, t_del = h.heapsort(d, n)
    assert t_ins > 0 and t_del > 0
    assert np.array_equal(d.to_numpy(np.uint32, n), np.sort(keys))


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("key,B", [("u32", 1024), ("u32", 4096), ("u32", 8192), ("u64", 1024), ("u64", 4096)])
@pytest.mark.parametrize("order", ["ascending", "descending", "ties"])
def test_heapsort_orders_large_batches(gpu, variant, key, B, order):
    # the batch sizes whose merges use group windows (copy windows in the
    # insertions, 64-lane groups for u32 up to 4096): keys in ascending order
    # (every insertion stops at once), descending (every window near the
    # root is a copy) and 12 distinct values (ties at every window edge)
    print(f"[case] heapsort orders {variant} {key} B={B} {order}", flush=True)
    n = 1 << 20
    keys = O.keys_u64(11, n) if key == "u64" else O.keys_u32(11, n)
    dt = keys.dtype.type
    if order == "ascending":
        keys = np.sort(keys)
    elif order == "descending":
        keys = np.sort(keys)[::-1].copy()
    else:
        keys = (keys % dt(12)) * dt(1000)
    h = Heap(batch_num_for(n, B), B, variant=variant, key=key)
    d = DeviceBuffer.from_numpy(keys)
    h.heapsort(d, n)
    assert h.error_flags() == 0 and h.isEmpty()
    assert np.array_equal(d.to_numpy(keys.dtype, n), np.sort(keys))
    h.close()


@pytest.mark.parametrize("variant", PB)
def test_partial_only_and_small_inputs(gpu, variant):
    # fewer keys than a batch: they live in the partial buffer only, and one
    # deleteRoot returns all of them (buitdd:305-320)
    B = 256
    for n in (1, 5, B - 1, B + 1, 3 * B + 17):
        keys = O.keys_u32(n, n)
        h = Heap(64, B, variant=variant)
        d = DeviceBuffer.from_numpy(keys)
        h.insert(d, n, chunk=B)
        assert h.checkInsertHeap(), h.last_check_message
        assert h.itemCount() == n
        out = DeviceBuffer(keys.nbytes)
        h.delete(out, n // B + 2, grid=4)
        G.sync()
        assert h.isEmpty()
        assert np.array_equal(out.to_numpy(np.uint32, n), np.sort(keys))


def test_empty_delete_and_errors(gpu):
    h = Heap(16, 256, variant="buitdd_np")
    out = DeviceBuffer(256 * 4)
    h.delete(out, 3)  # deleting from an empty heap returns nothing
    G.sync()
    assert h.isEmpty() and h.error_flags() == 0
    keys = DeviceBuffer.from_numpy(np.arange(100, dtype=np.uint32))
    with pytest.raises(G.BgpqError):
        h.insert(keys, 100, chunk=100)  # _np variants take full batches only
    with pytest.raises(G.BgpqError):
        Heap(16, 1000)  # batch size must be a power of two
    big = Heap(4, 256, variant="buitdd")
    with pytest.raises(G.BgpqError):
        big.heapsort(DeviceBuffer(4 * 256 * 4), 4 * 256)  # needs batch_num >= 2 * batches


def test_output_capacity_guard(gpu):
    """The heap-sort cursor never writes past the caller's output buffer: a
    deletion that would not fit is refused, flags error 16 and leaves its keys
    in the heap."""
    B = 256
    n = 8 * B
    keys = O.keys_u32(77, n)
    h = Heap(64, B, variant="buitdd_np")
    h.insert(DeviceBuffer.from_numpy(keys), n)
    small = DeviceBuffer(3 * B * 4)
    h.delete(small, 8, grid=4)
    G.sync()
    assert h.error_flags() & 16
    assert h.itemCount() == n - 3 * B
    srt = np.sort(keys)
    assert np.array_equal(small.to_numpy(np.uint32, 3 * B), srt[:3 * B])
    h.set_cursor(0)
    rest = DeviceBuffer((n - 3 * B) * 4)
    h.delete(rest, 8, grid=4)
    G.sync()
    assert h.isEmpty()
    assert np.array_equal(rest.to_numpy(np.uint32, n - 3 * B), srt[3 * B:])


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [64, 128, 256])
def test_concurrent_stress(gpu, variant, B):
    """Small batches and many workgroups maximise lock hand-offs: the
    concurrent harness on several seeds must still emit the sorted keys and
    leave the heap valid after the insertion phase."""
    for seed in range(4):
        print(f"[case] stress {variant} B={B} seed={seed}", flush=True)
        got, want, (ok, msg) = _harness(variant, B, 1 << 16, "u32", seed=100 + seed, grid=128)
        assert ok, msg
        assert np.array_equal(got, want), f"seed {seed}"


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [2048, 8192])
def test_concurrent_stress_large(gpu, variant, B):
    """The large-batch sift pipeline (half merges, locks requested ahead,
    children prefetched, the larger child released one level later) under
    concurrent insertions and deletions, on several seeds, once with keys
    folded onto 7 values so that merges and stop tests meet ties."""
    for seed, distinct in ((200, 0), (201, 0), (202, 7)):
        print(f"[case] stress large {variant} B={B} seed={seed} distinct={distinct}", flush=True)
        got, want, (ok, msg) = _harness(variant, B, 1 << 22, "u32", seed=seed, grid=64, distinct=distinct)
        assert ok, msg
        assert np.array_equal(got, want), f"seed {seed}"


@pytest.mark.parametrize("variant", ["buitdd_np", "buitdd"])
@pytest.mark.parametrize("B", [4096, 8192])
def test_early_take_repeat(gpu, variant, B):
    """Deleters of the bu variants take the last batch before the root lock.
    A deleter further ahead can still be refilling that batch's ancestor from
    its children, so the batch may hold keys below the root's largest; the
    root step then hands out the B smallest of both.  The window is narrow
    (about one run in a hundred of this harness met it), so the harness runs
    many times here, each output checked whole."""
    for it in range(40):
        seed = 1000 + it
        got, want, (ok, msg) = _harness(variant, B, 1 << 21, "u32", seed=seed, grid=32)
        assert ok, msg
        assert np.array_equal(got, want), f"seed {seed}: {int(np.count_nonzero(got != want))} keys out of place"
