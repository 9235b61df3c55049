"""GPU tests of the drop-in boundary as a caller binds it: mce_solve_csr on
caller-owned CSR arrays (the call an integration of BkSolverWrapper makes,
RDMCE/src/kernels/BkPivotBitBalance.cuh:611), and the rdmce CLI's stdout
protocol (RDMCE/src/mce_gpu.cu:20-21) as scripts/analyze.sh:97 parses it.
Integer results: bit-exact against the CPU oracle."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import rdmce_amd as R
from rdmce_amd import _lib
from oracle_lib import OracleGraph

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def solve_csr_ctypes(rowoff, col, labels, degeneracy):
    """mce_solve_csr through the raw C ABI: the arrays are numpy buffers the
    caller owns (u64 offsets, u32 columns and labels)."""
    L = _lib.lib()
    ro = np.ascontiguousarray(rowoff, dtype=np.uint64)
    co = np.ascontiguousarray(col, dtype=np.uint32)
    lb = np.ascontiguousarray(labels, dtype=np.uint32)
    opt = _lib.MceOptions()
    L.mce_options_init(ctypes.byref(opt))
    res = ctypes.POINTER(_lib.MceResult)()
    st = L.mce_solve_csr(ro.size - 1, ro.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                         co.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                         lb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), degeneracy, ctypes.byref(opt),
                         ctypes.byref(res))
    assert st == _lib.MCE_OK, L.mce_last_error()
    try:
        r = res.contents
        hist = np.ctypeslib.as_array(r.hist, shape=(r.hist_len,)).copy()
        return int(r.count), int(r.hash), {int(k): int(v) for k, v in enumerate(hist) if v}
    finally:
        L.mce_result_free(res)


@pytest.mark.parametrize("case", ["zachary", "er2048"])
def test_solve_csr_on_caller_arrays(case):
    banner(f"mce_solve_csr {case}")
    if case == "zachary":
        g = R.Graph()
        g.LoadFromFile(os.path.join(GOLDEN, "zachary.txt"))
    else:
        g = R.Graph.erdos_renyi(2048, 0.008, 1)
    g.SortByOrder(R.OrderType.DEG)
    ro = np.array(g.GetRowOffset(), dtype=np.uint64)
    co = np.array(g.GetColIdx(), dtype=np.uint32)
    lb = np.array(g.GetLabels(), dtype=np.uint32)
    count, h, hist = solve_csr_ctypes(ro, co, lb, g.GetDegeneracy())
    ref = OracleGraph.from_csr(ro, co, lb, g.GetDegeneracy()).mce()
    assert (count, h, hist) == (ref.count, ref.hash, ref.hist)
    if case == "zachary":
        assert count == 36 and hist == {2: 11, 3: 21, 4: 2, 5: 2}


def test_solve_csr_rejects_asymmetric_input():
    """An asymmetric caller CSR is an argument error, not a silent wrong
    count."""
    banner("mce_solve_csr asymmetric")
    L = _lib.lib()
    ro = np.array([0, 2, 3, 4], dtype=np.uint64)
    co = np.array([1, 2, 0, 1], dtype=np.uint32)
    opt = _lib.MceOptions()
    L.mce_options_init(ctypes.byref(opt))
    res = ctypes.POINTER(_lib.MceResult)()
    st = L.mce_solve_csr(3, ro.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                         co.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), None, 0, ctypes.byref(opt),
                         ctypes.byref(res))
    assert st == _lib.MCE_ERR_ARG and b"symmetric" in L.mce_last_error()


def test_cli_stdout_protocol(tmp_path):
    """rdmce -i g.txt prints the reference's result lines: 'MC numbers: N'
    and 'MCE GPU Time: <seconds>', whose 4th whitespace field is the number
    the reference's analysis script reads."""
    banner("rdmce CLI stdout")
    cli = os.path.join(ROOT, "rdmce_amd", "rdmce")
    for args in (["-i", os.path.join(GOLDEN, "zachary.txt")],
                 ["-i", os.path.join(GOLDEN, "zachary.txt"), "-o", "a", "-d", "0"]):
        r = subprocess.run([cli] + args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert re.search(r"^MC numbers: 36$", r.stdout, re.M), r.stdout
        line = next(x for x in r.stdout.splitlines() if x.startswith("MCE GPU Time:"))
        t = float(line.split()[3])
        assert 0 < t < 60
    # a preprocessed .bin input (convert mode, then solve) gives the same count
    src = tmp_path / "z.txt"
    src.write_text(open(os.path.join(GOLDEN, "zachary.txt")).read())
    assert subprocess.run([cli, "-i", str(src), "-c"], capture_output=True, timeout=120).returncode == 0
    r = subprocess.run([cli, "-i", str(tmp_path / "z.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and re.search(r"^MC numbers: 36$", r.stdout, re.M)
