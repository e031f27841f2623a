"""bench.py's contract that needs no GPU: it refuses more than one GPU
with a message instead of measuring, and reports profiled HBM traffic
only for the kernel sources the counters were collected for."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert r.stdout.strip() == "" or not r.stdout.lstrip().startswith("{")


def test_traffic_is_reported_only_for_the_profiled_sources(tmp_path):
    """roofline.traffic comes from a profile file that records the hash of
    the kernel sources its counters were collected for: a matching hash
    gives the figure, an edited source or a missing file gives None."""
    import hashlib
    import json
    sys.path.insert(0, ROOT)
    import bench
    src = tmp_path / "kernel.hip"
    src.write_text("// kernel v1\n")
    prof = tmp_path / "prof.json"
    digest = hashlib.sha256(src.read_bytes()).hexdigest()
    prof.write_text(json.dumps({"source_sha256": digest,
                                "hbm_bytes_per_launch": 1.5e8}))
    args = (str(prof), (str(src),), "hbm_bytes_per_launch")
    assert bench.profiled_traffic(*args) == 1.5e8
    src.write_text("// kernel v2\n")
    assert bench.profiled_traffic(*args) is None
    assert bench.profiled_traffic(str(tmp_path / "none.json"), (str(src),),
                                  "hbm_bytes_per_launch") is None


def test_matched_iterations_is_the_fewest_that_reach_the_target():
    sys.path.insert(0, ROOT)
    import bench
    # bad-1px after k = 0, 1, 2, ... refine iterations
    bad = [0.9, 0.30, 0.16, 0.14, 0.15]
    assert bench.matched_iterations(bad, 0.16) == 2
    assert bench.matched_iterations(bad, 0.145) == 3
    assert bench.matched_iterations(bad, 0.30) == 1
    assert bench.matched_iterations(bad, 0.10) is None
    # k = 0 (the proposal alone) never counts as a refine count
    assert bench.matched_iterations([0.05, 0.2], 0.1) is None


def test_fastest_matched_prefers_speed_among_matched_runs():
    sys.path.insert(0, ROOT)
    import bench
    runs = {"a": {"value": 10.0, "matched": True},
            "b": {"value": 30.0, "matched": False},
            "c": {"value": 20.0, "matched": True}}
    assert bench.fastest_matched(runs, "a") == "c"
    # an unmatched run never wins, however fast
    runs["c"]["matched"] = False
    assert bench.fastest_matched(runs, "x") == "a"
    runs["a"]["matched"] = False
    assert bench.fastest_matched(runs, "x") == "x"
