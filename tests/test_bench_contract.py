"""CPU checks of bench.py's contract that need no GPU: one GPU only, and the
HBM traffic figure of the headline roofline read from the committed
rocprofv3 counter summary."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_comes_from_the_committed_counter_summary():
    sys.path.insert(0, str(ROOT))
    import bench
    t = bench._traffic_from_profile("dgemm_nn_8192")
    # FETCH_SIZE + WRITE_SIZE of one 8192^3 launch: at least the algorithmic
    # bytes (A and B read once, C written once)
    assert t is not None and t >= 3 * 8192 * 8192 * 8
    assert bench._traffic_from_profile("no_such_workload") is None
