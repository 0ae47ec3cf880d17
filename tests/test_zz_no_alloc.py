"""The allocation-free lead-tape replay (espice_lanes_tab_refactor_solve_device)
against espice_lanes_tab_refactor_solve, bit for bit, on two calls in a row
(the second smaller), with device buffers the test owns."""
import numpy as np
import pytest

import egspice_amd as E
from test_gpu_parity import random_system


def say(msg):
    print(msg, flush=True)


def case(rng, n, lanes):
    while True:
        cp, ri, v = random_system(rng, n, 0.3)
        try:
            tab, idx = E.tab_build(n, cp, ri, v)
            break
        except E.EngineError:
            continue
    nnz = len(v)
    a = np.repeat(v[:, None], lanes, axis=1) * (0.95 + 0.1 * rng.random((nnz, lanes)))
    a[rng.integers(nnz), 0] = 0.0
    rhs = 2 * rng.random((n, lanes)) - 1
    return tab, idx, a, rhs


class Hip:
    """Device buffers through the HIP runtime the engine links (the test
    owns them; the entry point under test allocates nothing)."""

    def __init__(self):
        import ctypes as C
        import ctypes.util
        path = None
        for cand in ("/opt/rocm/lib/libamdhip64.so", ctypes.util.find_library("amdhip64")):
            if cand:
                try:
                    self.lib = C.CDLL(cand)
                    path = cand
                    break
                except OSError:
                    continue
        assert path, "libamdhip64 not found"
        self.C = C
        self.bufs = []

    def alloc(self, nbytes):
        C = self.C
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def put(self, arr):
        C = self.C
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.lib.hipMemcpy(C.c_void_p(p), arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
        return p

    def get(self, p, shape, dtype):
        C = self.C
        out = np.empty(shape, dtype=dtype)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
def test_no_alloc_replay_matches_the_allocating_one(engine):
    rng = np.random.default_rng(0x5C7A7C)
    hip = Hip()
    for n, lanes in ((24, 4096), (9, 300)):
        say(f"case: allocation-free replay n={n} lanes={lanes}")
        tab, idx, a, rhs = case(rng, n, lanes)
        dx_ref, fail_ref, _ = E.tab_refactor_solve(tab, idx, a, rhs)
        pa, pr = hip.put(a), hip.put(rhs)
        pdx = hip.put(np.full((n, lanes), np.nan))
        pf = hip.put(np.full(lanes, 7, dtype=np.uint8))
        E.tab_refactor_solve_device(tab, idx, lanes, pa, pr, pdx, pf)
        dx = hip.get(pdx, (n, lanes), np.float64)
        fail = hip.get(pf, (lanes,), np.uint8)
        hip.free()
        assert np.array_equal(fail, fail_ref)
        ok = fail == 0
        assert ok.sum() > 0
        assert np.array_equal(dx[:, ok].view(np.uint64), dx_ref[:, ok].view(np.uint64))


def test_no_alloc_rejects_what_does_not_fit(engine):
    """Sizes past the fixed scratch fail on the host before any GPU work."""
    rng = np.random.default_rng(3)
    tab, idx, a, rhs = case(rng, 40, 8)
    lanes = 1 << 24
    with pytest.raises(E.EngineError, match="ScratchTooSmall"):
        E.tab_refactor_solve_device(tab, idx, lanes, 1, 1, 1, 1)
