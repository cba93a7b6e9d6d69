"""bench.py contract on the CPU: one GPU only, byte formulas as documented, and the
measured-traffic lookup keyed on the exact workload."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_step_bytes_formulas():
    lds = {"lds_mode": True, "halo_entries": 100, "bond_slots": 0, "halo_max": 0}
    glob = {"lds_mode": False, "halo_entries": 0, "bond_slots": 0, "halo_max": 0}
    n_p, n_b = 128, 1000
    slices = 2
    assert bench.step_bytes(n_p, n_b, "fp64", lds)[0] == 10 * n_b + 168 * n_p + 8 * slices + 400
    assert bench.step_bytes(n_p, n_b, "mixed", lds)[0] == 2 * n_b + 200 * n_p + 8 * slices + 400
    assert bench.step_bytes(n_p, n_b, "fp64", glob)[0] == 12 * n_b + 168 * n_p + 8 * slices
    assert bench.step_bytes(n_p, n_b, "mixed", glob)[0] == 8 * n_b + 168 * n_p + 8 * slices
    assert bench.state_based_bytes("OSB", n_p, n_b, lds)[0] == 20 * n_b + 268 * n_p + 800
    assert bench.state_based_bytes("OSB", n_p, n_b, glob)[0] == 24 * n_b + 236 * n_p
    assert bench.state_based_bytes("OSB", n_p, n_b, lds, "mixed")[0] == 4 * n_b + 300 * n_p + 800
    assert bench.state_based_bytes("CPD", n_p, n_b, glob)[0] == 24 * n_b + 415 * n_p
    assert bench.state_based_bytes("CPD", n_p, n_b, lds)[0] == 20 * n_b + 511 * n_p + 800
    assert bench.reference_layout_bytes("OSB", n_p, n_b, "mixed")[0] == 22 * n_b + 248 * n_p
    assert bench.state_based_bytes("CPD", n_p, n_b, lds, "mixed")[0] == 4 * n_b + 414 * n_p + 800
    assert bench.state_based_bytes("CPD", n_p, n_b, lds, "fp64", "Wan")[0] == 20 * n_b + 607 * n_p + 800


def test_measured_traffic_lookup(tmp_path, monkeypatch):
    f = tmp_path / "t.md"
    f.write_text("# x\n`traffic fp64 points 10 bonds 20: 12345 bytes per launch`\n"
                 "traffic mixed points 10 bonds 20: 678\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic("fp64", 10, 20) == 12345
    assert bench.measured_traffic("mixed", 10, 20) == 678
    assert bench.measured_traffic("fp64", 11, 20) is None
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(tmp_path / "missing.md"))
    assert bench.measured_traffic("fp64", 10, 20) is None


def test_cpu_baseline_legs_on_small_samples():
    """The CPU-baseline legs run the oracle (its OpenMP loops) on the workloads of the GPU
    legs and report a positive rate with the cores used; small samples here."""
    sys.path.insert(0, os.path.join(ROOT, "peridynamics.jl_amd"))
    for kind in ("OSB", "CPD"):
        r = bench.cpu_state_based(kind, (12, 8, 6), target_seconds=0.2, max_steps=5)
        assert r["value"] > 0 and r["unit"] == "bond updates/s" and r["cores"] >= 1
        assert r["kind"] == "port" and kind in r["sample"]
    j = bench.cpu_jfnk((12, 12, 4), target_seconds=0.1)
    assert j["value"] > 0 and j["unit"] == "Krylov iterations/s" and j["ms_per_iteration"] > 0
    assert "GMRES(30)" in j["sample"]
