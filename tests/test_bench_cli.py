"""bench.py's contract parts that need no GPU: the single-GPU guard and the
HBM traffic figures it reads from the committed rocprofv3 summary."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_profile_parses_for_every_timed_kernel():
    import bench

    traffic = bench.measured_traffic()
    # the dominant kernel's HBM bytes per launch feed the JSON line's roofline.traffic
    for k in ("extend", "trace", "shade", "accumulate"):
        assert k in traffic and traffic[k] > 0, k
    assert all(isinstance(v, int) for v in traffic.values())
