"""bench.py's static inputs: the traffic figure it reports comes from the
per-shape table of the committed profile, and every launch shape has a
plausible row (at least the 16 bytes of output per ray of the 1080p frame)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_profiled_traffic_rows():
    sys.path.insert(0, ROOT)
    import bench
    import libsvo_amd as svo
    text = open(bench.PROFILE).read()
    table = text[text.index("## HBM traffic per launch shape"):]
    out_bytes = 1920 * 1080 * 16
    for shape in svo.LAUNCH_SHAPES:
        t = bench.profiled_traffic(shape)
        assert t is not None, shape
        assert out_bytes <= t < 2_000_000_000, (shape, t)
        row = next(line for line in table.splitlines() if line.startswith(f"| ({shape[0]}, {shape[1]}, {shape[2]}) |"))
        cells = [c.strip() for c in row.strip().strip("|").split("|")]
        assert cells[3] == f"{t:,}", (shape, row)
    assert bench.profiled_traffic((7, 7, "lds")) is None


def test_profiled_issue_rows():
    """Every launch shape has a row in the issue-counter table, with a VALU
    rate below the 2.0-per-CU-cycle ceiling and fractions in [0, 1]."""
    sys.path.insert(0, ROOT)
    import bench
    import libsvo_amd as svo
    for shape in svo.LAUNCH_SHAPES:
        r = bench.profiled_issue(shape)
        assert r is not None, shape
        assert 0.0 < r["achieved"] < r["peak"] == 2.0, (shape, r)
        assert 0.0 < r["simd_busy"] <= 1.0 and 0.0 < r["lane_use"] <= 1.0, (shape, r)
        assert abs(r["frac"] - r["achieved"] / 2.0) < 1e-3
    assert bench.profiled_issue((7, 7, "lds")) is None


def test_profiled_l2_hit_rows():
    """Every launch shape has an L2 hit rate in the traffic table, a fraction
    in (0, 1] (bench.py reports the tuned shape's as roofline.l2_hit_rate)."""
    sys.path.insert(0, ROOT)
    import bench
    import libsvo_amd as svo
    for shape in svo.LAUNCH_SHAPES:
        h = bench.profiled_l2_hit(shape)
        assert h is not None and 0.0 < h <= 1.0, (shape, h)
    assert bench.profiled_l2_hit((7, 7, "lds")) is None
