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
        assert row.rstrip(" |").endswith(f"{t:,}"), (shape, row)
    assert bench.profiled_traffic((7, 7, "lds")) is None
