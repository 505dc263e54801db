"""Host-side mirror of the reference's ``Heap<K>`` over the C ABI.

Names, argument meaning and error behaviour follow the reference class
(heap/buitdd/heap.cuh:13-231, heap/tditdd/heap.cuh:11-231) and its harness
kernels (test/heap_test.cu:13-74):

    Heap(batch_num, batch_size, init_limits, variant=..., key=...)
    reset(), checkInsertHeap(), itemCount(), nodeCount(), isEmpty(), printHeap()

plus the harness-level bulk operations the reference drives from its test
kernels: ``insert`` (insertKernel), ``delete`` (deleteKernel),
``concurrent`` (concurrentKernel) and ``heapsort``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

VARIANTS = {"buitdd": 0, "buitdd_np": 1, "tditdd": 2, "tditdd_np": 3}
KEYS = {"u32": (0, np.uint32), "u64": (1, np.uint64)}
CTR_BATCH_COUNT, CTR_PARTIAL, CTR_DELETE_COUNT, CTR_ERROR = 0, 1, 2, 3


class DeviceBuffer:
    """A raw device allocation owned by Python (bgpq_dev_malloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().bgpq_dev_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().bgpq_h2d(buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    def to_numpy(self, dtype, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().bgpq_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            check(lib().bgpq_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def free(self):
        if self.ptr:
            lib().bgpq_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync():
    check(lib().bgpq_sync())


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().bgpq_device_name(buf, 256))
    return buf.value.decode()


class Heap:
    """Heap<K>(batchNum, batchSize, init_limits) on the GPU.

    ``init_limits`` defaults to the key type's maximum, as the reference's
    harness passes UINT32_MAX (test/heap_test.cu:121).  Keys must be strictly
    below it.
    """

    def __init__(self, batch_num, batch_size, init_limits=None, variant="buitdd", key="u32"):
        if variant not in VARIANTS:
            raise ValueError(f"variant must be one of {sorted(VARIANTS)}")
        if key not in KEYS:
            raise ValueError(f"key must be one of {sorted(KEYS)}")
        self.variant = variant
        self.key = key
        self.key_type, self.dtype = KEYS[key]
        self.dtype = np.dtype(self.dtype)
        if init_limits is None:
            init_limits = int(np.iinfo(self.dtype).max)
        self.init_limits = int(init_limits)
        self.batch_num = int(batch_num)
        self.batch_size = int(batch_size)
        h = ctypes.c_void_p()
        check(lib().bgpq_create(ctypes.byref(h), VARIANTS[variant], self.key_type, self.batch_num,
                                self.batch_size, self.init_limits))
        self._h = h

    # ---- lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().bgpq_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self):
        check(lib().bgpq_reset(self._h, None))
        sync()

    # ---- host introspection (reference names)
    def checkInsertHeap(self):
        rc = check(lib().bgpq_check(self._h))
        self.last_check_message = "" if rc == 1 else _lib.last_error()
        return rc == 1

    def itemCount(self):
        v = ctypes.c_int64()
        check(lib().bgpq_item_count(self._h, ctypes.byref(v)))
        return v.value

    def nodeCount(self):
        v = ctypes.c_int64()
        check(lib().bgpq_node_count(self._h, ctypes.byref(v)))
        return v.value

    def isEmpty(self):
        return check(lib().bgpq_is_empty(self._h)) == 1

    def error_flags(self):
        v = ctypes.c_int()
        check(lib().bgpq_error_flags(self._h, ctypes.byref(v)))
        return v.value

    def snapshot(self):
        """(items, status, counters) as numpy arrays copied from the device."""
        items = np.empty((self.batch_num + 1) * self.batch_size, dtype=self.dtype)
        status = np.empty(self.batch_num + 1, dtype=np.int32)
        ctr = np.empty(8, dtype=np.int32)
        check(lib().bgpq_snapshot(self._h, items.ctypes.data,
                                  status.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  ctr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return items, status, ctr

    def view(self):
        """The heap's device storage (bgpq_view_t) for kernels written
        against include/bgpq_device.hpp."""
        v = _lib.View()
        check(lib().bgpq_view(self._h, ctypes.byref(v)))
        return v

    def vote_reset(self, busy=0):
        """Start a termination vote with `busy` seeding workgroups."""
        check(lib().bgpq_vote_reset(self._h, int(busy), None))

    def ifTerminate(self):
        """The termination vote has ended (reference: ifTerminate)."""
        return check(lib().bgpq_terminated(self._h)) == 1

    def clear(self):
        """Empty the heap, touching only the slots in use."""
        check(lib().bgpq_clear(self._h, None))
        sync()

    def printHeap(self):
        items, status, ctr = self.snapshot()
        b = self.batch_size
        print(f"batch partial {ctr[CTR_PARTIAL]}_{status[0]}:", *items[:ctr[CTR_PARTIAL]])
        for i in range(1, ctr[CTR_BATCH_COUNT] + 1):
            print(f"batch {i}_{status[i]}:", *items[i * b:(i + 1) * b])

    # ---- harness-level bulk operations
    def insert(self, d_keys, n, chunk=0, grid=0):
        check(lib().bgpq_insert(self._h, _ptr(d_keys), int(n), int(chunk), int(grid), None))

    def delete(self, d_out, calls, grid=0, out_keys=None):
        """deleteKernel: `calls` deletions writing at the heap-sort cursor;
        out_keys defaults to the capacity of d_out."""
        check(lib().bgpq_delete(self._h, _ptr(d_out), self._cap(d_out, out_keys), int(calls), int(grid), None))

    def concurrent(self, d_keys, n_insert, d_out, calls, chunk=0, grid=0, out_keys=None):
        check(lib().bgpq_concurrent(self._h, _ptr(d_keys), int(n_insert), int(chunk), _ptr(d_out),
                                    self._cap(d_out, out_keys), int(calls), int(grid), None))

    def set_cursor(self, value=0):
        check(lib().bgpq_set_cursor(self._h, int(value)))

    def _cap(self, d_out, out_keys):
        if out_keys is not None:
            return int(out_keys)
        if isinstance(d_out, DeviceBuffer):
            return d_out.nbytes // self.dtype.itemsize
        raise ValueError("out_keys is required for a raw device pointer")

    def heapsort(self, d_keys, n, chunk=0, grid_insert=0, grid_delete=0, d_out=None):
        """Heap-sort n device keys into d_out (default: in place); returns
        (ms_insert, ms_delete) measured with HIP events."""
        a, b = ctypes.c_float(), ctypes.c_float()
        out = d_keys if d_out is None else d_out
        check(lib().bgpq_heapsort(self._h, _ptr(d_keys), _ptr(out), int(n), int(chunk), int(grid_insert),
                                  int(grid_delete), ctypes.byref(a), ctypes.byref(b), None))
        return a.value, b.value


def _ptr(buf):
    if buf is None:
        return None
    if isinstance(buf, DeviceBuffer):
        return buf.ptr
    return int(buf)


def batch_num_for(n_keys, batch_size):
    """The reference harness's sizing: the smallest power of two with
    batchNum * batchSize >= 2 * n (test/heap_test.cu:97-100)."""
    bn = 1
    while bn * batch_size < 2 * n_keys:
        bn <<= 1
    return bn
