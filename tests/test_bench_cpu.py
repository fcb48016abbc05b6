"""bench.py's contract on the CPU: one GPU only, and the committed HBM
traffic summary is the one measured with the benchmark's default launch
configuration (so the default run reports `roofline.traffic`)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""


def test_traffic_summary_matches_the_default_configuration():
    a = bench.arg_parser().parse_args([])
    traffic = bench.measured_traffic(a.exchange_mode, a.kernel_variant, a.wg, bool(a.fused_x), a.wg_alpha)
    assert traffic is not None
    assert (a.exchange_mode, a.kernel_variant, a.wg, a.fused_x, a.wg_alpha) == (1, 35, 256, 1, 512)
    algorithmic = 152 * 256 ** 3
    assert algorithmic <= traffic < 1.05 * algorithmic


def test_traffic_summary_is_not_used_for_another_configuration():
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=0, wg=256, fused_x=True, wg_alpha=512) is None
    assert bench.measured_traffic(exchange_mode=0, kernel_variant=35, wg=256, fused_x=True, wg_alpha=512) is None
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=35, wg=128, fused_x=True, wg_alpha=512) is None
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=35, wg=256, fused_x=False) is None
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=35, wg=256, fused_x=True) is None   # alpha at 256
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=35, wg=256, fused_x=True, wg_alpha=256) is None
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=3, wg=256, fused_x=True, wg_alpha=512) is None   # eighths
    assert bench.measured_traffic(exchange_mode=1, kernel_variant=7, wg=256, fused_x=True, wg_alpha=512) is None   # no map