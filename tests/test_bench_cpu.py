"""The bench line's roofline: its formulas and the measured traffic it reads
from the committed counter summary (no GPU needed)."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import bench  # noqa: E402


class _Sites:
    n_atoms = 12288


D = {"S": _Sites, "K": (50, 50, 50), "n_pairs": 1_320_000}


def test_measured_traffic_rows_for_headline_kernels():
    for kname in ("k_plane_matvec", "k_spread_cw"):
        traffic, source = bench.measured_traffic(kname)
        assert traffic and traffic > 0, kname
        assert source.startswith(bench.COUNTER_SUMMARY)
    assert bench.measured_traffic("no_such_kernel") == (None, None)


def test_roofline_plane_path_and_fused_bytes():
    prof = {"fft": (0.6, 48), "influence": (0.1, 4), "plane_matvec": (0.6, 24)}
    fft = bench.roofline("fft", prof["fft"], D, "mixed", prof=prof)
    assert fft["kernel"] == "k_fz_r2c/k_fz_c2r"          # 2 z launches per plane launch
    pm = bench.roofline("plane_matvec", prof["plane_matvec"], D, "mixed", prof=prof)
    G, Gh, N, P, w = 125000, 50 * 50 * 26, 12288, 1_320_000, 8
    assert pm["algorithmic_bytes_per_launch"] == 4 * Gh * w + 8 * Gh + 2 * P * (4 + 2 * w) + N * 3 * (w + 8)
    assert abs(pm["frac"] - pm["achieved"] / bench.HBM_PEAK_GBS) < 1e-12
    line = {"fft": (0.6, 96), "influence": (0.1, 24)}      # 4 launches per x pass: line-major path
    assert bench.roofline("fft", line["fft"], D, "mixed", prof=line)["kernel"] == "k_fz_r2c/k_fy/k_fz_c2r"
