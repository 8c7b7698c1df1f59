"""bench.py reads the HBM traffic and the vector-memory pipeline figures of
its roofline from the committed counter summaries: both parse, and agree
with the figures the summaries state."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_profiled_figures_parse():
    import bench
    for wl in ("sphere_field", "glass_torus"):
        for k in ("kExtend", "kShadow"):
            t = bench.profiled_traffic(wl, k)
            p = bench.profiled_pipeline(wl, k)
            assert t is not None and t > 0, (wl, k)
            assert p is not None, (wl, k)
            assert 0.0 < p["ta_busy"] <= 1.0 and 0.0 < p["td_busy"] <= 1.0, (wl, k, p)
            assert 0.0 < p["l1_lookups_per_cu_cycle"] <= 1.05, (wl, k, p)
    # no summary for the box room: the line reports no traffic rather than a guess
    assert bench.profiled_traffic("box_room", "kExtend") is None
    assert bench.profiled_pipeline("box_room", "kExtend") is None


def test_multi_gpu_request_exits_with_message():
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
