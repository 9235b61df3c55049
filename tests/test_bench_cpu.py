"""The bench's host-side pieces on the CPU: the roofline arithmetic (three
ceilings, the binding one named, the shared per-root denominator and the
figures reported beside it), the committed build-term fixture, and the
refusal of more than one GPU."""
import json
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_roofline_ceilings_and_binding_bound():
    out = types.SimpleNamespace(build_bytes=1_000_000_000_000, search_bytes=8 * 10**12)
    plain = types.SimpleNamespace(build_bytes=16 * 10**12)
    r = bench.roofline(out, 2000.0, "no-such-workload", plain)
    hbm = 16e12 / (bench.HBM_PEAK_GBS * 1e9)
    lds = 8e12 / (bench.LDS_PEAK_GBS * 1e9)
    valu = bench.OPS_PER_WORD * 1e12 / (bench.VALU_PEAK_GOPS * 1e9)
    assert r["ceilings_ms"] == pytest.approx({"hbm": hbm * 1e3, "lds": lds * 1e3, "valu": valu * 1e3})
    assert r["bound"] == max(("hbm", hbm), ("lds", lds), ("valu", valu), key=lambda kv: kv[1])[0]
    assert r["frac"] == pytest.approx(max(hbm, lds, valu) / 2.0)
    assert r["build_bytes"] == plain.build_bytes
    assert r["own_build"]["build_bytes"] == out.build_bytes
    assert r["own_build"]["frac"] == pytest.approx(max(1e12 / (bench.HBM_PEAK_GBS * 1e9), lds, valu) / 2.0)
    assert r["per_root_build_formula"] is None


def test_roofline_without_plain_leg_uses_own_bytes():
    out = types.SimpleNamespace(build_bytes=4 * 10**12, search_bytes=8)
    r = bench.roofline(out, 1000.0, "no-such-workload")
    assert r["bound"] == "hbm" and r["build_bytes"] == out.build_bytes
    assert r["frac"] == pytest.approx(4e12 / (bench.HBM_PEAK_GBS * 1e9) / 1.0)


def test_build_terms_fixture_is_consistent():
    with open(os.path.join(ROOT, "tests", "golden", "build_terms.json")) as f:
        terms = json.load(f)
    assert {"rmat24-core500", "cl4m-core300", "rmat20-core499"} <= set(terms)
    for t in terms.values():
        assert t["build_bytes"] == 4 * t["deg"] + 4 * t["pdeg"] + 8 * t["rows_w"]
    assert terms["rmat24-core500"]["roots"] == 210705255
    f = bench.per_root_build("rmat24-core500", 2.0)
    assert f["hbm_ms"] == pytest.approx(terms["rmat24-core500"]["build_bytes"] / (bench.HBM_PEAK_GBS * 1e6))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
