"""MI355X-native batched GPU priority queue (BGPQ).

The product is the HIP library gpuheap_amd/lib/libbgpq.so (C ABI in
include/bgpq.h); this package is its Python host mirror.
"""
from .heap import DeviceBuffer, Heap, batch_num_for, device_name, sync, VARIANTS, KEYS  # noqa: F401
from ._lib import BgpqError, LIB_PATH, lib  # noqa: F401
