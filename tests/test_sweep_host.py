"""CPU checks of the multi-scene sweep's host logic (pyroclastmpm_amd.batch):
the scene parameters drawn from the seed and the column geometry every scene
of the batched and differentiated forms is built from. No GPU is touched."""
import math

import numpy as np

from pyroclastmpm_amd.batch import column_geometry, sweep_parameters


def test_sweep_parameters_ranges_and_seed():
    p = sweep_parameters(64, seed=1)
    assert len(p) == 64
    n, h, phi, E = (np.array(v) for v in zip(*p))
    assert n.min() >= 16000 and n.max() <= 1000000
    assert h.min() >= 0.01 and h.max() <= 0.02
    assert phi.min() >= 20.0 and phi.max() <= 40.0
    assert E.min() >= 5e5 and E.max() <= 2e6
    # log-uniform counts: the logs spread over the whole range
    ln = np.log(n)
    assert ln.min() < math.log(16000) + 0.5 and ln.max() > math.log(1000000) - 0.5
    assert sweep_parameters(64, seed=1) == p
    assert sweep_parameters(64, seed=2) != p


def test_column_geometry_count_and_bounds():
    for n_target, h in [(20000, 0.02), (150000, 0.013)]:
        pos, end = column_geometry(n_target, h, per_axis=2, jitter=0.1, seed=3)
        assert pos.dtype == np.float32 and pos.shape[1] == 3
        # lattice rounding: within a few percent of the target
        assert abs(pos.shape[0] - n_target) <= 0.06 * n_target, (pos.shape[0], n_target)
        # the column is 0.2 x 0.5 x 0.4 of the domain's 1.0 x 0.5 x 0.5 scale
        s = end[0]
        assert np.isclose(end[1], 0.5 * s) and np.isclose(end[2], 0.5 * s)
        spacing = h / 2
        lo = -0.1 * spacing
        assert pos.min() >= lo
        assert pos[:, 0].max() <= 0.2 * s + 0.1 * spacing
        assert pos[:, 1].max() <= 0.5 * s + 0.1 * spacing
        assert pos[:, 2].max() <= 0.4 * s + 0.1 * spacing
        # same seed, same positions
        assert np.array_equal(column_geometry(n_target, h, 2, 0.1, 3)[0], pos)
