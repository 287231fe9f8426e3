"""The parts of bench.py's contract that need no GPU: one GPU only (--gpus N > 1 exits with
a message and status 2), the HBM traffic figures the roofline objects report are read from
the committed rocprofv3 summaries, and the synchronisation helper degrades to a no-op."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_figures_come_from_the_committed_profiles():
    for path in (bench.C1_TRAFFIC_PROFILE, bench.C2_TRAFFIC_PROFILE):
        assert os.path.exists(path)
        t = bench.measured_traffic(path)
        assert t is not None and t > 0
        # the figure is FETCH_SIZE (copy-calibrated) + WRITE_SIZE of the summary's table, in KiB
        rows = {}
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 3 and cells[0] in ("FETCH_SIZE", "WRITE_SIZE"):
                try:
                    rows[cells[0]] = float(cells[2])
                except ValueError:
                    pass
        assert abs(t - 1024 * (rows["FETCH_SIZE"] + rows["WRITE_SIZE"])) < 1e-3 * t
    assert bench.measured_traffic(os.path.join(HERE, "no_such_profile.md")) is None


def test_device_sync_without_a_gpu():
    sync = bench.device_sync()
    assert callable(sync)
    sync()
