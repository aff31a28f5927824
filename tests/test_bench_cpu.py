"""bench.py's contract on the CPU: more than one GPU is refused with a
message, and the algorithmic byte counts behind `roofline.achieved` are the
documented ones."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_per_particle_step():
    # 1M cylinder: d = 3, S = 3, E = 6 -> 288 B; 2D dam break: d = 2, S = 1, E = 4 -> 192 B
    assert bench.algorithmic_bytes_per_particle_step(3, 3, 6) == 288
    assert bench.algorithmic_bytes_per_particle_step(2, 1, 4) == 192


def test_ssa_bytes_per_event():
    # the cylinder's A+B->C: 124 B per reaction, 320 B per diffusion (6 neighbours)
    assert bench.ssa_bytes_per_event(3, 1, 6) == (124, 320)
    roof = bench.ssa_roofline(3, 1, 6, 1000, 500, 10.0, "k_ssa")
    assert roof["events"] == 1500
    assert abs(roof["achieved"] - (124 * 1000 + 320 * 500) / 1e-2 / 1e9) < 1e-3
    assert roof["bound"] == "hbm" and roof["peak"] == bench.HBM_PEAK_GBS
