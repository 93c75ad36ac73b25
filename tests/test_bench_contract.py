"""bench.py's command-line contract on the CPU (no GPU work is started)."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_profiles_resolve_by_kernel_name():
    """The committed per-launch traffic of the stage kernel is found under
    the name the bench line reports for the tile the autotuner picked."""
    for prec, ft, tiles in (("mixed", "float", [(32, 8)]), ("fp64", "double", [(32, 16)])):
        t = bench.load_traffic("stage_kernel_%s_riemann4q_64x64_blocks_of_64sq_traffic.json" % prec)
        assert t is not None
        for ti, tj in tiles:
            k = t["kernels"]["stage_kernel<%s,%d,%d,fused>" % (ft, ti, tj)]
            assert abs(k["fetch_bytes_per_launch"] + k["write_bytes_per_launch"] - k["bytes_per_launch"]) < 1.0
            # at least the algorithmic 162 B per cell-stage of the 16.8 M-cell workload
            assert k["bytes_per_launch"] > bench.BYTES_PER_CELL_STAGE * 64 * 64 * 64 * 64
    assert json.load(open(os.path.join(ROOT, "profiles", "stage_kernel_mixed_riemann4q_64x64_blocks_of_64sq_traffic.json")))
