"""bench.py's command line on the CPU: the one-GPU contract and the committed
traffic figures its roofline line reports (no GPU is touched)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""  # no JSON line for a configuration that was not measured


def test_headline_traffic_figures_are_committed():
    """The headline's dominant kernels have HBM traffic per launch from the
    counters, within 1 % of their algorithmic bytes (1024^3, R=40)."""
    import bench
    n = 1024 ** 3
    for precision, kernel, bytes_per_voxel in (("fp32_split", "xypass", 4 + 0.125 + 4),
                                               ("fp32_split", "zpass", 4 + 4),
                                               ("bf16", "xypass", 2 + 0.125 + 2),
                                               ("bf16", "zpass", 2 + 2)):
        got = bench.load_traffic("c2", precision, kernel)
        assert got is not None, (precision, kernel)
        assert abs(got / (n * bytes_per_voxel) - 1) < 0.01, (precision, kernel, got)
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        entries = json.load(f)
    assert all(abs(v["ratio"] - 1) < 0.02 for k, v in entries.items() if k.startswith("c2/")), entries
