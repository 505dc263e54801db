"""Pins the CPU oracle (oracle/bgpq_model.c, oracle/keys.c) before it is
trusted as the checker of the GPU heap.

* heap-sort output equals numpy's sort (the reference checks against
  std::sort, test/heap_test.cu:118,165-170);
* every deleteRoot returns exactly the smallest keys held, checked against
  a brute-force multiset on tiny batch sizes;
* the restated checkInsertHeap holds after every operation and catches
  planted corruptions;
* the bit-reversed slot map matches a string reversal and is a tree;
* SplitMix64 matches its published first outputs.
"""
import collections

import numpy as np
import pytest

import _oracle as O

VARIANTS = list(O.VARIANTS)
PB = ("buitdd", "tditdd")


def test_splitmix64_known_answer():
    # First outputs of SplitMix64 from state 0 (Vigna's reference splitmix64.c).
    assert O.splitmix64_stream(0, 2) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]


def test_key_ranges():
    k = O.keys_u32(0, 1 << 16)
    assert k.max() <= 0xFFFFFFFE
    k64 = O.keys_u64(0, 1 << 12)
    assert np.array_equal(k64 & np.uint64(0xFFFFFFFF), np.arange(1 << 12, dtype=np.uint64))
    assert np.array_equal((k64 >> np.uint64(32)).astype(np.uint32), O.keys_u32(0, 1 << 12))


def test_rev_idx_brute_force():
    for x in list(range(1, 1 << 14)) + [(1 << 31) + 12345, (1 << 30) + 7, 0x7FFFFFFF]:
        b = bin(x)[3:]  # bits below the leading one
        want = int("1" + b[::-1], 2)
        assert O.rev_idx(x) == want, x
    assert O.rev_idx(1) == 1


@pytest.mark.parametrize("variant", VARIANTS)
def test_parent_child_tree(variant):
    # every slot 2..N has a parent in the level above, and each parent has
    # exactly two children (slots of one level map onto the level above)
    for lvl in range(1, 12):
        kids = collections.Counter(O.parent(variant, i) for i in range(1 << lvl, 1 << (lvl + 1)))
        assert set(kids) == set(range(1 << (lvl - 1), 1 << lvl))
        assert set(kids.values()) == {2}


def _keys(kind, n, seed):
    if kind == "random":
        return O.keys_u32(seed, n).astype(np.uint64)
    if kind == "descending":
        return np.sort(O.keys_u32(seed, n)).astype(np.uint64)[::-1].copy()
    if kind == "ascending":
        return np.sort(O.keys_u32(seed, n)).astype(np.uint64)
    if kind == "dups":
        return (O.keys_u32(seed, n) % 7).astype(np.uint64)
    raise ValueError(kind)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [4, 8, 64])
@pytest.mark.parametrize("kind", ["random", "descending", "ascending", "dups"])
def test_model_heapsort_equals_sort(variant, B, kind):
    for n_batches in (1, 3, 37):
        n = n_batches * B
        keys = _keys(kind, n, seed=n_batches)
        bn = 4 * n_batches + 4
        out = O.heapsort(variant, B, bn, O.U32_MAX, keys, chunk=B)
        assert np.array_equal(out, np.sort(keys))


@pytest.mark.parametrize("variant", PB)
@pytest.mark.parametrize("B", [8, 64, 256])
def test_model_partial_chunks(variant, B):
    # the partial-buffer harness inserts B/3- and 2B/3-key calls
    # (test/partial_heap_test.cu:17,49)
    for chunk in (max(1, B // 3), max(1, 2 * B // 3), 1, B - 1):
        for n in (chunk, 5 * B + 3, 17 * B):
            keys = O.keys_u32(n + chunk, n).astype(np.uint64)
            bn = 4 * (n // B + 2)
            out = O.heapsort(variant, B, bn, O.U32_MAX, keys, chunk=chunk)
            assert np.array_equal(out, np.sort(keys))


def _brute_force_run(variant, B, seed, steps=400, key_mod=None):
    rng = np.random.default_rng(seed)
    m = O.Model(variant, 512, B, O.U32_MAX)
    held = collections.Counter()
    pb = variant in PB
    for _ in range(steps):
        total = sum(held.values())
        if total == 0 or rng.random() < 0.55:
            size = int(rng.integers(1, B + 1)) if pb else B
            if total + size > 200 * B:
                continue
            keys = rng.integers(0, 2**32 - 1, size=size, dtype=np.uint64)
            if key_mod:
                keys %= np.uint64(key_mod)
            assert m.insert(keys) == 0
            held.update(int(k) for k in keys)
        else:
            before = m.st.delete_count
            size = m.delete()
            got = m.out[before:before + size]
            want = sorted(held.elements())[:size]
            # a delete returns a whole batch, or the partial buffer when no batch exists
            assert size == (B if total >= B else total)
            assert list(map(int, got)) == want
            held.subtract(int(k) for k in got)
            held += collections.Counter()
        ok, why = m.check()
        assert ok, why
        bc, pbs, _, err = m.counters()
        assert err == 0
        assert bc * B + pbs == sum(held.values())
        # the held multiset equals the keys in the heap's live slots
        items = m.items()
        live = list(items[:pbs]) + list(items[B:(bc + 1) * B])
        assert collections.Counter(int(k) for k in live) == held


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("B", [4, 8])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_model_brute_force_priority_queue(variant, B, seed):
    _brute_force_run(variant, B, seed)


@pytest.mark.parametrize("variant", VARIANTS)
def test_model_brute_force_duplicates(variant):
    _brute_force_run(variant, 4, 11, key_mod=3)


@pytest.mark.parametrize("variant", VARIANTS)
def test_checker_catches_planted_corruption(variant):
    B = 8
    m = O.Model(variant, 64, B, O.U32_MAX)
    keys = O.keys_u32(5, 10 * B).astype(np.uint64)
    for i in range(0, len(keys), B):
        m.insert(keys[i:i + B])
    assert m.check()[0]
    items = np.ctypeslib.as_array(m.st.items, shape=((m.batch_num + 1) * B,))
    # plant: break ascending order inside batch 3
    items[3 * B], items[3 * B + 1] = items[3 * B + 1] + 1, items[3 * B]
    ok, why = m.check()
    assert not ok and "ascending" in why
    items[3 * B], items[3 * B + 1] = items[3 * B + 1], items[3 * B] - 1
    assert m.check()[0]
    # plant: a child batch below its parent's last key
    child = 5
    p = O.parent(variant, child)
    saved = items[child * B]
    items[child * B] = items[p * B + B - 1] - 1 if items[p * B + B - 1] > 0 else 0
    if items[child * B] < items[p * B + B - 1]:
        ok, why = m.check()
        assert not ok
    items[child * B] = saved
    assert m.check()[0]


def _early_take_script(check):
    """Three deleters A, A2, C of a buitdd_np heap (B = 4), interleaved as
    they can be on the GPU, where deleteRoot takes the last batch before it
    waits for the root:

      A  pops the root, takes slot 7 and sifts: its root step descends into
         slot 3 (the node it then holds);
      A2 takes slot 6, then pops the root A refilled;
      C  takes slot 5, a child of slot 3, while A has not yet refilled
         slot 3 from its children;
      A  refills slot 3 without slot 5's keys, A2 refills the root;
      C  pops the root: without the root check it would hand out a key
         above the smallest one it carries (24 < 50).

    Slots hold logical nodes rev(slot): 5 <-> 6 and 6 <-> 5.  Returns the
    model, the concatenated output and whether C's root check merged."""
    m = O.Model("buitdd_np", 16, 4, O.U32_MAX)
    tree = {1: [0, 1, 2, 3], 2: [10, 11, 12, 90], 3: [20, 21, 22, 23], 4: [91, 92, 93, 94],
            6: [95, 96, 97, 98],  # logical node 5, a child of slot 2
            5: [24, 25, 26, 27],  # logical node 6, a child of slot 3
            7: [50, 51, 52, 53]}
    for slot, keys in tree.items():
        m.set_items(slot, keys)
    m.st.batch_count = 7
    assert m.check()[0]
    # A: deleteRoot + deleteUpdate's root step
    cand_a = m.take_last()
    m.root_out()
    cur_a = m.sift_step(1, cand_a)
    assert cur_a == 3
    # A2: early take, then the root
    cand_a2 = m.take_last()
    assert not m.root_out(cand_a2, check)
    # C: early take of slot 5, a child of the node A holds
    assert O.parent("buitdd_np", m.st.batch_count) == cur_a
    cand_c = m.take_last()
    # A finishes at slot 3, then A2 sifts from the root
    while cur_a:
        cur_a = m.sift_step(cur_a, cand_a)
    cur = 1
    while cur:
        cur = m.sift_step(cur, cand_a2)
    assert m.check()[0]  # consistent without slot 5's keys
    merged = m.root_out(cand_c, check)
    cur = 1
    while cur:
        cur = m.sift_step(cur, cand_c)
    assert m.check()[0]
    while m.delete():
        assert m.check()[0]
    return m, m.out[:m.st.delete_count].copy(), merged


def test_early_take_root_check_restores_order():
    m, out, merged = _early_take_script(check=True)
    assert merged
    # C handed out the B smallest of root and candidate: 24 joined 21..23
    assert list(out[8:12]) == [21, 22, 23, 24]
    all_keys = list(range(4)) + [10, 11, 12, 90, 20, 21, 22, 23, 91, 92, 93, 94, 95, 96, 97, 98, 24, 25, 26, 27,
                                 50, 51, 52, 53]
    assert list(out) == sorted(all_keys)
    assert m.st.error == 0


def test_early_take_without_root_check_misorders():
    # the same interleaving with the check left out: C hands out 50 before 24
    _, out, merged = _early_take_script(check=False)
    assert not merged
    assert list(out[8:12]) == [21, 22, 23, 50]
    assert not np.all(out[1:] >= out[:-1])


@pytest.mark.parametrize("variant", PB)
def test_model_overflow_leaves_partial_buffer(variant):
    # a full heap: case 2.2 needs a new slot; nothing may change
    B = 8
    m = O.Model(variant, 2, B, O.U32_MAX)
    assert m.insert(np.arange(100, 100 + B, dtype=np.uint64)) == 0
    assert m.insert(np.arange(200, 200 + B, dtype=np.uint64)) == 0
    assert m.insert(np.arange(300, 303, dtype=np.uint64)) == 0  # partial buffer: 3 keys
    before = m.items(), m.counters()
    assert m.insert(np.arange(400, 400 + B, dtype=np.uint64)) == -1
    after = m.items(), m.counters()
    assert np.array_equal(before[0], after[0])
    assert after[1][:3] == before[1][:3] and after[1][3] == 1
