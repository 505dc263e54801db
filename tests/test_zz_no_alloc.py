"""bgpq_knapsack_noalloc (include/bgpq_apps.h) through the C ABI: two calls in
a row, the second on a smaller instance, each with the same optimum as
bgpq_knapsack on the same heap and as the CPU dynamic program; an instance
whose dominance cells do not fit the scratch array is refused with a clear
error and changes nothing."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpuheap_amd import Heap
from gpuheap_amd import apps as A
from gpuheap_amd._lib import AppStats, BgpqError, check, last_error, lib

pytestmark = pytest.mark.gpu


def _noalloc(heap, w, v, cap, grid=8):
    w = np.ascontiguousarray(w, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    best = ctypes.c_uint64()
    st = AppStats()
    rc = lib().bgpq_knapsack_noalloc(heap._h, len(w), w.ctypes.data_as(ctypes.c_void_p),
                                     v.ctypes.data_as(ctypes.c_void_p), cap, 0, grid, ctypes.byref(best),
                                     ctypes.byref(st), None)
    return rc, best.value, st.as_dict()


def test_knapsack_noalloc_matches_allocating_entry_point(gpu):
    print("[case] knapsack_noalloc: n=40 then n=24, against bgpq_knapsack and the DP", flush=True)
    with A.app_heap(batch_size=256, batch_num=1 << 14) as h:
        for n, seed in ((40, 0), (24, 1)):
            w, v, cap = A.knapsack_instance(n, seed)
            rc, got, st = _noalloc(h, w, v, cap)
            assert rc == 0, last_error()
            assert st["error_flags"] == 0 and h.isEmpty()
            want, st_alloc = A.knapsack(h, w, v, cap, grid=8)
            assert got == want == O.knapsack_dp(w, v, cap)
            assert st["popped"] > 0 and st["pushed"] > 0


def test_knapsack_noalloc_refuses_what_does_not_fit(gpu):
    print("[case] knapsack_noalloc: n=200 does not fit the scratch array", flush=True)
    with A.app_heap(batch_size=256, batch_num=1 << 14) as h:
        w, v, cap = A.knapsack_instance(200, 2)
        rc, _, _ = _noalloc(h, w, v, cap)
        assert rc < 0 and "scratch array" in last_error()
        assert h.isEmpty()
        # and a call after the refusal still solves
        w, v, cap = A.knapsack_instance(30, 3)
        rc, got, _ = _noalloc(h, w, v, cap)
        assert rc == 0 and got == O.knapsack_dp(w, v, cap)
