"""The committed HBM-traffic profile that bench.py reads for its roofline.

bench.py reports `traffic` for whatever tile shapes its autotuning run keeps,
so the profile of the default batch size must hold a figure for every
(direction, level, shape) the autotuner can choose; an average over the
shapes a run kept is then always available.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def test_traffic_profile_covers_every_shape():
    batch = 32
    assert os.path.exists(bench.traffic_profile(batch))
    for direction in ("forward", "inverse"):
        for shape in range(5):
            t = bench.measured_traffic(direction, [shape] * 5, batch)
            assert t is not None, (direction, shape)
            # At least the compulsory bytes of one launch, at most a few times them.
            algorithmic = bench.dwt_algorithmic_bytes(True) * batch / bench.LEVELS
            assert algorithmic <= t < 4 * algorithmic, (direction, shape, t)
    # A mix per level, as an autotuned run keeps.
    assert bench.measured_traffic("inverse", [4, 4, 4, 2, 1], batch) is not None
    # No profile for another batch size: no figure.
    assert bench.measured_traffic("inverse", [0] * 5, 7) is None
