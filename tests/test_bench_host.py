"""Host-side parts of bench.py (CPU only): the one-GPU contract and the
pure-Python reference subsample leg."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode != 0
    assert "--gpus 1" in proc.stderr
    assert proc.stdout.strip() == ""  # no JSON line for a run that was not made


def test_python_subsample_leg_reports_rates_and_labelled_extrapolations():
    sys.path.insert(0, ROOT)
    import bench

    out = bench.python_subsample_leg()
    ray = out["ray_anyhit_2M_triangles_1M_rays"]
    knn = out["knn16_4M_x_4M_mixture"]
    assert ray["rays_per_s_at_sample"] > 0 and ray["extrapolated_full_s_upper_bound"] > 0
    assert knn["queries_per_s_at_sample"] > 0 and knn["extrapolated_full_s"] > 0
    assert out["pip_positive_1000_polygons"]["points_per_s"] > 0
    frn = out["fixed_radius_r0.01_1e5_points"]
    assert frn["extrapolated_queries_per_s_at_1e7_points"] < frn["queries_per_s"]
    for v in out.values():
        assert "sample" in v
