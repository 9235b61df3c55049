"""The HIP kernels under host emulation (tests/emu): compiled for the CPU
against an emulated HIP runtime, one OS thread per lane, device memory and
LDS ending at guard pages, every cross-lane operation checked to be reached
by all lanes of its wave.  Runs a quick subset of the emulation parity cases
(tests/emu/run_cases.py) against the CPU oracle; no GPU needed."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
LIB = os.path.join(EMU, "_build", "librdmce_emu.so")

QUICK = ["zachary", "closed K5", "closed moon_moser4", "closed cocktail8", "closed multipartite_1_5_9_2",
         "G(60,0.5) tau=1 lds=1", "G(60,0.5) tau=-1 lds=0", "G(120,0.3) tau=4 lds=0", "G(200,0.08) tau=24 lds=1",
         "vertex roots", "two-word subset", "deferred pass", "batch", "plan graph replay",
         "pack zachary list", "pack G(60,0.5) slots=1", "pack G(45,0.8) slots=4", "pack wide runs K90",
         "pack wide runs K150", "pack routed runs", "pack batch",
         "choose root mode"]


@pytest.fixture(scope="module")
def emu_lib():
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("host clang++ not available")
    r = subprocess.run(["make", "-C", EMU, "-s"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    return LIB


@pytest.mark.parametrize("case", QUICK)
def test_emulated_kernels_match_oracle(emu_lib, case):
    env = dict(os.environ, RDMCE_LIB=emu_lib, EMU_CUS="1", EMU_BARRIER_TIMEOUT="120")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_cases.py"), case], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ok" in r.stdout and "FAIL" not in r.stdout


@pytest.mark.parametrize("case", ["pack G(60,0.5) slots=1", "pack G(45,0.8) slots=4", "pack wide runs",
                                  "pack routed runs"])
def test_emulated_packed_depth_first_expansion(emu_lib, case):
    """The packed search's bounded-stack mode (only the top record expanded)
    on every dispatch beyond two stacked records, and a pool with room for
    two kept vertex bitmaps (the other wide roots are built on their own by
    the second launch): same results."""
    lib = os.path.join(EMU, "_build", "librdmce_emu_thr.so")
    env = dict(os.environ, RDMCE_LIB=lib, EMU_CUS="1", EMU_BARRIER_TIMEOUT="120")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_cases.py"), case], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ok" in r.stdout and "FAIL" not in r.stdout
