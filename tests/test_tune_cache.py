"""Autotune cache file handling (include/svo_hip.h, svo_tune_cache_*): no
GPU needed.  The cache has no counterpart in the reference, whose launch
shape is the fixed work size of its renderToBuffer enqueue
(projects/raycast_serialized_opencl/src/raycast_serialized_opencl.cpp:729-749)."""
import os

import pytest

import libsvo_amd as svo
from libsvo_amd import _capi

KEY_A = "arch=gfx950;cus=256;policy=0;views=1;rays=2097152;scenes=13/10/4096/full"
KEY_B = "arch=gfx950;cus=256;policy=2;views=64;rays=67108864;scenes=13/10/4096/full,13/9/4096/full"


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_missing_file_is_an_empty_cache(tmp_path):
    c = svo.TuneCache(str(tmp_path / "none.tsv"))
    assert len(c) == 0


def test_round_trip(tmp_path):
    p = str(tmp_path / "cache.tsv")
    _write(p, f"# comment\n\n{KEY_A}\t16\t1\tlds\t0.27761\n{KEY_B}\t8\t2\tlds_node\t2.96\n")
    c = svo.TuneCache(p)
    assert len(c) == 2
    q = str(tmp_path / "copy.tsv")
    c.save(q)
    lines = [l for l in open(q).read().splitlines() if l and not l.startswith("#")]
    assert sorted(lines) == sorted([f"{KEY_A}\t16\t1\tlds\t0.27761", f"{KEY_B}\t8\t2\tlds_node\t2.96000"])
    assert not os.path.exists(q + ".tmp")
    d = svo.TuneCache(q)
    assert len(d) == 2
    # a later file overrides an entry with the same key, others are kept
    _write(p, f"{KEY_A}\t32\t4\tprivate\t0.5\n")
    d.load(p)
    assert len(d) == 2
    d.save(q)
    assert f"{KEY_A}\t32\t4\tprivate\t0.50000" in open(q).read()


@pytest.mark.parametrize("line", [
    f"{KEY_A} 16 1 lds 0.2",             # no tabs
    f"{KEY_A}\t16\t1\tlds",              # no time
    f"{KEY_A}\t12\t1\tlds\t0.2",         # no such tile width
    f"{KEY_A}\t16\t3\tlds\t0.2",         # no such wave count
    f"{KEY_A}\t16\t1\tglobal\t0.2",      # no such stack layout
])
def test_malformed_line_is_refused_and_nothing_merged(tmp_path, line):
    p = str(tmp_path / "bad.tsv")
    _write(p, f"{KEY_B}\t8\t1\tlds\t1.0\n{line}\n")
    c = svo.TuneCache()
    with pytest.raises(svo.SvoError) as e:
        c.load(p)
    assert e.value.status == _capi.SVO_ERR_FORMAT
    assert "bad.tsv:2" in str(e.value)
    assert len(c) == 0


def test_save_to_missing_directory_fails(tmp_path):
    c = svo.TuneCache()
    with pytest.raises(svo.SvoError) as e:
        c.save(str(tmp_path / "no" / "such" / "dir.tsv"))
    assert e.value.status == _capi.SVO_ERR_IO
    with pytest.raises(ValueError):
        c.save()
