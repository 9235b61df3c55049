"""The allocation-free entry point mce_solve_device_csr_noalloc (all scratch
in a static device array) against the entry point it varies,
mce_solve_csr, through the C ABI: bit-exact count, clique hash and size
histogram, and the same search-tree node count, on two calls in a row (the
second with a smaller graph), plus the size check that rejects graphs
beyond the array."""
import ctypes
import os
import sys

import numpy as np
import pytest

import rdmce_amd as R
from rdmce_amd import _lib
from oracle_lib import splitmix64

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HIST = 64


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def lib():
    L = _lib.lib()
    f = L.mce_solve_device_csr_noalloc
    f.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int32
    return L


def reference(g):
    """mce_solve_csr on the same CSR (host arrays)."""
    L = _lib.lib()
    ro = np.ascontiguousarray(g.GetRowOffset(), dtype=np.uint64)
    co = np.ascontiguousarray(g.GetColIdx(), dtype=np.uint32)
    lb = np.ascontiguousarray(g.GetLabels(), dtype=np.uint32)
    opt = _lib.MceOptions()
    L.mce_options_init(ctypes.byref(opt))
    res = ctypes.POINTER(_lib.MceResult)()
    st = L.mce_solve_csr(ro.size - 1, ro.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                         co.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                         lb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), g.GetDegeneracy(), ctypes.byref(opt),
                         ctypes.byref(res))
    assert st == _lib.MCE_OK, L.mce_last_error()
    try:
        r = res.contents
        hist = np.ctypeslib.as_array(r.hist, shape=(r.hist_len,)).copy()
        return int(r.count), int(r.hash), int(r.total_nodes), hist
    finally:
        L.mce_result_free(res)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime librdmce.so itself
    loaded (test-side device buffers; the entry point under test allocates
    nothing)."""

    def __init__(self):
        _lib.lib()
        path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln)
        self.h = ctypes.CDLL(path)
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(arr.nbytes, 8)) == 0
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def get(self, p, count, dtype):
        out = np.zeros(count, dtype=dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def run_noalloc(hip, g):
    ro = np.ascontiguousarray(g.GetRowOffset(), dtype=np.uint64)
    co = np.ascontiguousarray(g.GetColIdx(), dtype=np.uint32)
    lh = np.array([splitmix64(int(x)) for x in g.GetLabels()], dtype=np.uint64)
    d_ro, d_co, d_lh = hip.put(ro), hip.put(co), hip.put(lh)
    d_out = hip.put(np.full(4 + HIST, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64))
    st = lib().mce_solve_device_csr_noalloc(0, ro.size - 1, co.size, d_ro, d_co, d_lh, HIST, None, d_out)
    assert st == _lib.MCE_OK, _lib.lib().mce_last_error()
    return hip.get(d_out, 4 + HIST, np.uint64)


def test_noalloc_matches_solve_csr_twice_second_smaller():
    first = R.Graph.erdos_renyi(2048, 0.008, 1)
    first.SortByOrder(R.OrderType.DEG)
    second = R.Graph()
    second.LoadFromFile(os.path.join(GOLDEN, "zachary.txt"))
    second.SortByOrder(R.OrderType.DEG)
    dense = R.Graph.from_edges([(u, v) for u in range(90) for v in range(u + 1, 90) if (u * 7 + v * 13) % 5])
    dense.SortByOrder(R.OrderType.DEG)
    hip = Hip()
    try:
        for name, g in (("ER G(2048, 0.008)", first), ("zachary (smaller)", second), ("dense 90", dense),
                        ("zachary again", second)):
            banner(f"noalloc {name}")
            out = run_noalloc(hip, g)
            count, h, nodes, hist = reference(g)
            assert int(out[3]) == 0, hex(int(out[3]))
            assert (int(out[0]), int(out[1]), int(out[2])) == (count, h, nodes), name
            k = min(HIST, hist.size)
            assert np.array_equal(out[4:4 + k], hist[:k]), name
            assert int(out[4:].sum()) == count
        assert int(run_noalloc(hip, second)[0]) == 36
    finally:
        hip.free()


def test_noalloc_rejects_graphs_beyond_the_array():
    banner("noalloc size limit")
    hip = Hip()
    try:
        p = hip.put(np.zeros(16, dtype=np.uint64))
        L = lib()
        st = L.mce_solve_device_csr_noalloc(0, 10_000_000, 0, p, p, p, HIST, None, p)
        assert st == _lib.MCE_ERR_UNSUPPORTED
        assert b"no-alloc scratch" in L.mce_last_error()
        assert L.mce_solve_device_csr_noalloc(0, 10, 3, p, p, p, HIST, None, p) == _lib.MCE_ERR_ARG
    finally:
        hip.free()
