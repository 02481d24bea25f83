"""bench.py measures one GPU: asking for more exits with status 2 and says so."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_profile_figures_parse():
    """bench.py reads the traffic, L2 hit rate and VALU issue / lane figures of each workload's
    committed rocprofv3 summary; every summary names the command that made it."""
    sys.path.insert(0, ROOT)
    import bench

    for name in ("holes", "ico"):
        f = bench.profile_figures(name)
        assert set(f) == {"traffic_bytes_per_walk_step", "l2_hit_rate", "valu_issue_share", "valu_lane_activity",
                          "valu_issue_rate_frac"}, name
        assert f["traffic_bytes_per_walk_step"] > 0 and 0 < f["l2_hit_rate"] <= 1
        assert 0 < f["valu_issue_share"] <= 1 and 0 < f["valu_lane_activity"] <= 1
        assert 0 < f["valu_issue_rate_frac"] < 2
        text = open(os.path.join(ROOT, "profiles", f"k_fused_pers_{name}.md")).read()
        assert "rocprofv3 --kernel-trace --stats" in text and "rocprofv3 --pmc FETCH_SIZE" in text
