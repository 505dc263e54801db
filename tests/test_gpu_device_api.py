"""The public device API and the checker, on the GPU.

* A heap consumer built outside the library (tests/ext/ext_knapsack.hip:
  its own persistent kernel over include/bgpq_device.hpp and the C ABI,
  linked with libbgpq.so) runs the reference's delete -> expand -> insert
  loop (artifacts/knapsack/models_pq.cuh:50-156) with the threshold
  deleteRoot of the A* forks and the termination vote; its optimum equals
  the CPU dynamic program.
* checkInsertHeap (bgpq_check) reports each class of planted corruption,
  with the same first violation as the oracle's orc_check on the same
  state (heap/buitdd/heap.cuh:118-172).
* An insertion that overflows the heap changes nothing, the partial
  buffer included, and flags error 1.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import gpuheap_amd as G
from gpuheap_amd import DeviceBuffer, Heap
from gpuheap_amd import apps as A
from gpuheap_amd._lib import check, lib

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ext", "_build", "ext_knapsack")


@pytest.mark.parametrize("n,seed,grid", [(24, 3, 1), (32, 3, 8), (40, 3, 16), (40, 5, 64), (36, 9, 32)])
def test_external_kernel_knapsack(gpu, tmp_path, n, seed, grid):
    print(f"[case] external knapsack kernel n={n} seed={seed} grid={grid}", flush=True)
    assert os.path.exists(EXE), "tests/ext/_build/ext_knapsack is not built (make -C tests/ext)"
    w, v, cap = A.knapsack_instance(n, seed)
    inst = tmp_path / "inst.data"
    # the reference's format: "n capacity", then "benefit weight" per item
    inst.write_text(f"{n} {cap}\n" + "".join(f"{int(b)} {int(a)}\n" for a, b in zip(w, v)))
    r = subprocess.run([EXE, str(inst), str(grid)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["best"] == O.knapsack_dp(w, v, cap)
    assert res["terminated"] == 1 and res["empty_after_clear"] == 1 and res["heap_flags"] == 0
    assert res["popped"] > 0


def _write(ptr, offset_bytes, arr):
    arr = np.ascontiguousarray(arr)
    check(lib().bgpq_h2d(ctypes.c_void_p(ptr + offset_bytes), arr.ctypes.data, arr.nbytes))


def _read(ptr, offset_bytes, dtype, count):
    out = np.empty(count, dtype=dtype)
    check(lib().bgpq_d2h(out.ctypes.data, ctypes.c_void_p(ptr + offset_bytes), out.nbytes))
    return out


def _built_pair(variant, B=64, batches=10, partial=20):
    """The same state in a GPU heap (one workgroup) and in the model."""
    h = Heap(64, B, variant=variant)
    m = O.Model(variant, 64, B, O.U32_MAX)
    keys = O.keys_u32(17, batches * B + partial)
    d = DeviceBuffer.from_numpy(keys)
    n_full = batches * B
    h.insert(d, n_full, chunk=B, grid=1)
    for i in range(0, n_full, B):
        m.insert(keys[i:i + B].astype(np.uint64))
    if partial and variant in ("buitdd", "tditdd"):
        d2 = DeviceBuffer.from_numpy(keys[n_full:])
        h.insert(d2, partial, chunk=partial, grid=1)
        m.insert(keys[n_full:].astype(np.uint64))
    G.sync()
    assert h.checkInsertHeap() and m.check()[0]
    return h, m


def _model_items(m):
    return np.ctypeslib.as_array(m.st.items, shape=((m.batch_num + 1) * m.B,))


@pytest.mark.parametrize("variant", list(G.VARIANTS))
def test_checker_reports_planted_corruption(gpu, variant):
    print(f"[case] checker corruption {variant}", flush=True)
    B = 64
    h, m = _built_pair(variant, B)
    v = h.view()
    mi = _model_items(m)

    def both(slot_key, value):
        """Set key `slot_key` to `value` in both heaps; return the old value."""
        old = int(mi[slot_key])
        mi[slot_key] = value
        _write(v.items, 4 * slot_key, np.array([value], np.uint32))
        return old

    def outcomes():
        ok_g = h.checkInsertHeap()
        ok_m, why_m = m.check()
        return ok_g, h.last_check_message, ok_m, why_m

    cases = []
    # 1. a batch not ascending: swap two distinct keys of batch 3
    a, b = 3 * B + 5, 3 * B + 6
    if mi[a] != mi[b]:
        cases.append(("not ascending", [(a, int(mi[b])), (b, int(mi[a]))]))
    # 2. a child below its parent's last key (the child stays ascending)
    child = 5
    p = O.parent(variant, child)
    cases.append(("below parent", [(child * B, int(mi[p * B + B - 1]) - 1)]))
    if variant in ("buitdd", "tditdd"):
        # 3. the partial buffer below the root's last key
        cases.append(("partial buffer below the root", [(0, int(mi[2 * B - 1]) - 1)]))
        # 4. the partial buffer not ascending
        if mi[3] != mi[4]:
            cases.append(("partial buffer not ascending", [(3, int(mi[4])), (4, int(mi[3]))]))
    for name, edits in cases:
        olds = [both(k, val) for k, val in edits]
        ok_g, why_g, ok_m, why_m = outcomes()
        assert not ok_g and not ok_m, name
        assert why_g == why_m, (name, why_g, why_m)
        for (k, _), old in zip(reversed(edits), reversed(olds)):
            both(k, old)
        assert h.checkInsertHeap() and m.check()[0], name
    # 5. a status word that is not AVAIL (the model keeps no status words)
    for slot, want in ((2, "state of batch 2 is 1, expected AVAIL"), (0, "partial batch state 1, expected AVAIL")):
        off = 4 * slot * v.word_stride
        _write(v.status, off, np.array([1], np.int32))
        assert not h.checkInsertHeap()
        assert h.last_check_message == want
        _write(v.status, off, np.array([0], np.int32))
        assert h.checkInsertHeap()


@pytest.mark.parametrize("variant", ["buitdd", "tditdd"])
def test_overflow_keeps_partial_buffer(gpu, variant):
    print(f"[case] overflow {variant}", flush=True)
    B = 64
    h = Heap(2, B, variant=variant)
    m = O.Model(variant, 2, B, O.U32_MAX)
    for lo, n in ((100, B), (1000, B), (5000, 3)):
        keys = np.arange(lo, lo + n, dtype=np.uint32)
        h.insert(DeviceBuffer.from_numpy(keys), n, chunk=n, grid=1)
        assert m.insert(keys.astype(np.uint64)) == 0
    G.sync()
    before, _, ctr0 = h.snapshot()
    # a full batch needs a third slot: nothing may change
    keys = np.arange(9000, 9000 + B, dtype=np.uint32)
    h.insert(DeviceBuffer.from_numpy(keys), B, chunk=B, grid=1)
    assert m.insert(keys.astype(np.uint64)) == -1
    G.sync()
    after, status, ctr = h.snapshot()
    assert np.array_equal(before, after)
    assert np.array_equal(after.astype(np.uint64), m.items())
    assert tuple(ctr[:3]) == tuple(ctr0[:3]) == (2, 3, 0)
    assert ctr[3] == 1 and m.st.error == 1
    assert not status.any()


def test_view_matches_handle(gpu):
    h = Heap(32, 256, variant="tditdd_np", key="u64")
    v = h.view()
    assert (v.batch_num, v.batch_size, v.variant, v.key_type) == (32, 256, 3, 1)
    assert v.word_stride == 32 and v.init_limits == O.U64_MAX
    # the view's items are the heap's: every slot holds init_limits
    assert np.all(_read(v.items, 0, np.uint64, 33 * 256) == np.uint64(O.U64_MAX))
