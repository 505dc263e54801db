"""The reference's own harness flow (test/heap_test.cu, test/partial_heap_test.cu)
written in C++ over include/bgpq_heap.hpp (tests/cpp/heap_test.cpp): insert
half, insert the other half while deleting as many, delete the rest, output
in place equal to std::sort; checkInsertHeap holds after the first phase."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "heap_test")


@pytest.mark.parametrize("B,logn,variant,chunk", [
    (256, 18, "buitdd", 85), (512, 19, "buitdd_np", 512), (1024, 20, "tditdd", 341),
    (1024, 20, "tditdd_np", 1024), (2048, 20, "buitdd", 1365), (8192, 21, "buitdd_np", 8192),
])
def test_cpp_heap_harness(gpu, B, logn, variant, chunk):
    print(f"[case] cpp harness B={B} 2^{logn} {variant} chunk={chunk}", flush=True)
    assert os.path.exists(EXE), "tests/cpp/_build/heap_test is not built (make -C tests/cpp)"
    r = subprocess.run([EXE, str(B), str(logn), variant, str(chunk), "7"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "Success" in r.stdout, r.stdout + r.stderr


def test_cpp_print_heap(gpu):
    """printHeap of the C++ facade prints the reference's lines: the partial
    buffer as "batch partial <size>_<state>:", then "batch <slot>_<state>:"
    per batch in use, each with its keys; here keys 0..199 at B = 64."""
    print("[case] cpp printHeap", flush=True)
    assert os.path.exists(EXE), "tests/cpp/_build/heap_test is not built (make -C tests/cpp)"
    r = subprocess.run([EXE, "print"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("batch")]
    heads = [ln.split(":")[0] for ln in lines]
    assert heads == ["batch partial 8_0", "batch 1_0", "batch 2_0", "batch 3_0"], heads
    rows = [[int(x) for x in ln.split(":")[1].split()] for ln in lines]
    assert [len(r) for r in rows] == [8, 64, 64, 64]
    assert all(r == sorted(r) for r in rows)
    assert sorted(sum(rows, [])) == list(range(200))
    assert rows[1][0] == 0  # the root holds the smallest keys
