"""Runs the batch primitives alone, for counter collection under rocprofv3:
the group merge, the two half merges of a sift step and the low-register merge
path (bgpq_test_merge algo 2, 4, 5), then the pad-and-sort of the insert path
(bgpq_test_sort), each on NB random u32 batches, one workgroup per batch.

    python -m gpuheap_amd.primbench [B] [NB]
"""
import sys

import numpy as np

from . import DeviceBuffer, sync
from ._lib import check, lib


def main(argv):
    B = int(argv[1]) if len(argv) > 1 else 8192
    nb = int(argv[2]) if len(argv) > 2 else 512
    rng = np.random.default_rng(1)
    a = np.sort(rng.integers(0, 2**32 - 1, size=(nb, B), dtype=np.uint32), axis=1)
    b = np.sort(rng.integers(0, 2**32 - 1, size=(nb, B), dtype=np.uint32), axis=1)
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    small, large = DeviceBuffer(a.nbytes), DeviceBuffer(a.nbytes)
    for algo in (2, 4, 5):
        check(lib().bgpq_test_merge(0, B, da.ptr, db.ptr, small.ptr, large.ptr, nb, algo))
        sync()
        u = np.sort(np.concatenate([a[0], b[0]]))
        if not (np.array_equal(small.to_numpy(np.uint32, B), u[:B])):
            raise SystemExit(f"merge algo {algo}: wrong result")
    sizes = DeviceBuffer.from_numpy(np.full(nb, B, dtype=np.int32))
    check(lib().bgpq_test_sort(0, B, da.ptr, sizes.ptr, small.ptr, 0xFFFFFFFF, nb))
    sync()
    print(f"primbench: B={B}, {nb} batches: merges (algo 2, 4, 5) and sort done", flush=True)


if __name__ == "__main__":
    main(sys.argv)
