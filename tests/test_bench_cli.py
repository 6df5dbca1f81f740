"""bench.py's command line on the CPU: every workload has its handler and
the single-GPU contract (--gpus above 1 exits with a message)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_every_workload_has_a_handler():
    import bench
    for fn in ("run_fixed", "oracle_fixed", "run_disk", "oracle_disk", "run_apply", "run_nodes", "main"):
        assert callable(getattr(bench, fn)), fn
    assert {"square_4k_n21", "square_1m_n35", "square_1m_n56", "cube_2m_n80", "disk_2m_uneven",
            "apply_1m_n35", "nodes_1m_n35"} <= set(bench.WORKLOADS)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
