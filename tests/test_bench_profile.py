"""The committed HBM-traffic profile that bench.py reads for its roofline.

bench.py reports `traffic` for whatever tile shapes its autotuning run keeps,
so the profile of the default batch size must hold a figure for every
(direction, level, shape) the autotuner can choose; an average over the
shapes a run kept is then always available.  The fused schedules
(each component's first two levels in one launch, luma level 2 alone) have
their own figures for the fused launch and the level-2 shapes the profiled
run kept, in both directions.
"""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def test_traffic_profile_covers_every_shape():
    batch = 32
    assert os.path.exists(bench.traffic_profile(batch))
    for direction in ("forward", "inverse"):
        for shape in range(5):
            t = bench.measured_traffic(direction, [shape] * 5, batch, False)
            assert t is not None, (direction, shape)
            # At least the compulsory bytes of one launch, at most a few times them.
            algorithmic = bench.dwt_algorithmic_bytes(True) * batch / bench.LEVELS
            assert algorithmic <= t < 4 * algorithmic, (direction, shape, t)
    # A mix per level, as an autotuned run keeps.
    assert bench.measured_traffic("inverse", [4, 4, 4, 2, 1], batch, False) is not None
    # No profile for another batch size: no figure.
    assert bench.measured_traffic("inverse", [0] * 5, 7, False) is None


@pytest.mark.parametrize("direction", ["forward", "inverse"])
def test_traffic_profile_fused(direction):
    batch = 32
    assert (bench.fused_forward if direction == "forward" else bench.fused_inverse)(batch)
    with open(bench.traffic_profile(batch)) as f:
        lines = [l.split() for l in f]
    assert any(p[:2] == ["traffic_fused", direction] for p in lines)
    kept = [int(p[3]) for p in lines if p[:3] == ["traffic_fused_level", direction, "2"]]
    assert kept
    for shape in kept:
        t = bench.measured_traffic(direction, [0, 0, shape, shape, shape], batch, True)
        algorithmic = bench.dwt_algorithmic_bytes(True) * batch / (bench.LEVELS - 1)
        assert t is not None and algorithmic <= t < 4 * algorithmic, (shape, t)


def test_fused_inverse_tiles():
    # 120-sample-wide tiles over the 3840x2160 luma plane and two 1920x1080
    # chroma planes: 32 x 34 + 2 x 16 x 17 (64 rows), 32 x 68 + 2 x 16 x 34 (32 rows).
    assert bench.inv_fused_tiles(1, 64) == 1632
    assert bench.inv_fused_tiles(1, 32) == 3264
    assert bench.inv_fused_tile_h(32) == 64 and bench.inv_fused_tile_h(8) == 32 and bench.inv_fused_tile_h(1) == 16
