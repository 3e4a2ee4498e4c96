"""The on-device reduction's order, restated in numpy (tests/test_gpu_output.py
compares the GPU with it bit for bit), pinned by exact sums and extrema."""
import math

import numpy as np

from test_gpu_output import fold


def test_fold_restatement_pins():
    rng = np.random.default_rng(0)
    for N in (1, 2, 255, 256, 257, 4099, 100003):
        x = rng.standard_normal(N) * 10.0 ** rng.uniform(-3, 3, N)
        assert abs(fold(x, "sum") - math.fsum(x)) <= 1e-13 * np.abs(x).sum()
        assert fold(x, "min") == x.min() and fold(x, "max") == x.max()
        assert fold(x, "mean") == fold(x, "sum") / N
        # small integers add exactly in any order
        k = rng.integers(-1000, 1000, N).astype(np.float64)
        assert fold(k, "sum") == k.sum()
    # NaN ignored by min / max, like fmin / fmax
    x = np.array([3.0, np.nan, -1.0, 2.0])
    assert fold(x, "min") == -1.0 and fold(x, "max") == 3.0
