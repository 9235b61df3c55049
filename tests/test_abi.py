"""The C-ABI library loads and exports every function include/rdmce.h
declares; argument errors return status codes (no exit).  CPU only: no
compute call touches a GPU here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import rdmce_amd as R
from rdmce_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rdmce.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mce_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["mce_solve", "mce_solve_csr", "mce_graph_load_coo_text", "mce_graph_load_bin",
                 "mce_graph_store_bin", "mce_graph_order", "mce_result_free", "mce_plan_create", "mce_plan_run"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mce_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_cli_builds_and_rejects_bad_usage():
    cli = os.path.join(ROOT, "rdmce_amd", "rdmce")
    assert os.path.exists(cli)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([cli, "-i", "x.txt", "-d", "0,1"], capture_output=True, text=True)
    assert r.returncode == 1


def test_cli_convert_only(tmp_path):
    cli = os.path.join(ROOT, "rdmce_amd", "rdmce")
    src = tmp_path / "z.txt"
    src.write_text(open(os.path.join(ROOT, "tests", "golden", "zachary.txt")).read())
    r = subprocess.run([cli, "-i", str(src), "-c"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "vertices: 34 \tedges: 78" in r.stdout
    assert "max degree: 17 \tdegeneracy: 4" in r.stdout
    g = R.Graph()
    g.LoadFromBin(str(tmp_path / "z.bin"))
    assert g.GetNumEdges() == 78


def test_argument_errors_are_status_codes():
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.mce_graph_load_coo_text(None, ctypes.byref(out)) == _lib.MCE_ERR_ARG
    assert b"null" in L.mce_last_error()
    assert L.mce_gen_er(10, 1.5, 0, ctypes.byref(out)) == _lib.MCE_ERR_ARG
    assert L.mce_gen_rmat(0, 16, 0.5, 0.2, 0.2, 0, ctypes.byref(out)) == _lib.MCE_ERR_ARG


def test_solve_without_device_fails_cleanly():
    if R.device_count() > 0:
        pytest.skip("a GPU is present")
    g = R.Graph.from_edges([(0, 1), (1, 2)])
    with pytest.raises(R.MceError) as e:
        R.solve(g)
    assert e.value.status == _lib.MCE_ERR_DEVICE


def test_options_struct_roundtrip():
    o = _lib.MceOptions()
    _lib.lib().mce_options_init(ctypes.byref(o))
    assert o.root_mode == _lib.ROOT_EDGE and o.lds_tier == 1 and o.schedule == 0
    assert o.word_path == 0  # register paths on by default


@pytest.mark.parametrize("cname,mirror", [("mce_options", "MceOptions"), ("mce_result", "MceResult"),
                                           ("mce_tune_result", "MceTuneResult"),
                                           ("mce_batch_result", "MceBatchResult")])
def test_struct_layout_matches_header(tmp_path, cname, mirror):
    """The ctypes mirrors of the ABI's structs have the C structs' sizes and
    field offsets (compiled from include/rdmce.h with the host compiler)."""
    cls = getattr(_lib, mirror)
    fields = [f for f, _ in cls._fields_ if f != "reserved"]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "rdmce.h"\nint main(void) {\n'
                   f'  printf("size %zu\\n", sizeof({cname}));\n'
                   + "".join(f'  printf("{f} %zu\\n", offsetof({cname}, {f}));\n' for f in fields)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                           text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(cls)
    for f in fields:
        assert int(out[f]) == getattr(cls, f).offset, f"{cname}.{f} offset"



def _csr_status(rowoff, col):
    L = _lib.lib()
    ro = np.ascontiguousarray(rowoff, dtype=np.uint64)
    co = np.ascontiguousarray(col, dtype=np.uint32)
    out = ctypes.c_void_p()
    st = L.mce_graph_from_csr(ro.size - 1, ro.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              co.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), None, 0, ctypes.byref(out))
    if st == 0:
        L.mce_graph_free(out)
    return st, L.mce_last_error()


def test_graph_from_csr_rejects_asymmetric_and_offset_errors():
    """mce_graph_from_csr (the input of mce_solve_csr) accepts a symmetric
    CSR and rejects one that lists an edge in one direction only, a first
    offset other than 0, self loops and unsorted rows: the kernels build P
    from out-lists and X from in-lists, so such input would change results
    without an error."""
    # path 0-1-2, symmetric
    assert _csr_status([0, 1, 3, 4], [1, 0, 2, 1])[0] == _lib.MCE_OK
    # 0 lists 1 and 2, but 2 does not list 0 (even entry count, sorted rows)
    st, msg = _csr_status([0, 2, 3, 4], [1, 2, 0, 1])
    assert st == _lib.MCE_ERR_ARG and b"symmetric" in msg
    # same adjacency behind a non-zero first offset
    st, msg = _csr_status([2, 3, 5, 6], [9, 9, 1, 0, 2, 1])
    assert st == _lib.MCE_ERR_ARG and b"rowoff[0]" in msg
    st, msg = _csr_status([0, 1, 2], [0, 1])
    assert st == _lib.MCE_ERR_ARG and b"self loop" in msg
    st, msg = _csr_status([0, 2, 3, 4], [2, 1, 0, 0])
    assert st == _lib.MCE_ERR_ARG and b"ascending" in msg


def test_graph_from_csr_roundtrip_of_a_loaded_graph():
    g = R.Graph()
    g.LoadFromFile(os.path.join(ROOT, "tests", "golden", "zachary.txt"))
    g.SortByOrder(R.OrderType.DEG)
    h = R.Graph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    assert np.array_equal(h.GetRowOffset(), g.GetRowOffset())
    assert np.array_equal(h.GetColIdx(), g.GetColIdx())
    assert h.GetNumEdges() == 78
