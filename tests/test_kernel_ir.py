"""Static checks of the HIP kernels' optimized device code for gfx950 (CPU
only: hipcc cross-compiles).

A private array that the code indexes with a run-time value either stays in
scratch memory (an `alloca` in address space 5 of the optimized IR) or is
promoted to a vector read and written at a run-time position
(extractelement / insertelement with a non-constant index, lowered to
v_movrel or index-mode moves).  The kernels are written so that neither
occurs: every per-lane array is indexed by constants or through uniform
selects.  This is the check the host emulation cannot make (on the CPU a
register array is ordinary memory)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rdmce_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

DYNAMIC_VEC = re.compile(r"(extractelement <[^>]+> [^,]+, i(32|64) %|insertelement <[^>]+> [^,]+, [^,]+, i(32|64) %)")


@pytest.mark.parametrize("src", ["mce_kernel.hip", "pack_kernel.hip", "noalloc.cpp", "mlp.hip"])
def test_no_private_arrays_indexed_at_run_time(tmp_path, src):
    if not os.path.exists(HIPCC) or shutil.which("hipcc") is None and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path / (src + ".ll")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-emit-llvm",
                        "-S", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC,
                        "-o", str(out)] + lang + [os.path.join(CSRC, src)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    ir = out.read_text()
    allocas = [ln for ln in ir.splitlines() if re.search(r"=\s*alloca\b", ln)]
    assert not allocas, "private arrays left in scratch:\n" + "\n".join(allocas[:10])
    dyn = [ln.strip() for ln in ir.splitlines() if DYNAMIC_VEC.search(ln)]
    assert not dyn, "vectors indexed at run time:\n" + "\n".join(dyn[:10])
