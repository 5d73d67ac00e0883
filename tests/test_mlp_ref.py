"""Pins the CPU reference of the MLP gradient (tests/mlp_ref.py) by central
finite differences in float64: the check the device gradients are then held
to. No GPU."""
import numpy as np

import mlp_ref


def test_reference_gradient_is_the_derivative():
    rng = np.random.default_rng(42)
    p = rng.normal(0, 0.2, 5313)
    p[12 * 64:16 * 64] = 0.0
    x = np.zeros((48, 16))
    x[:, :12] = rng.uniform(0, 1.2, (48, 12))
    t = rng.uniform(1, 10, 48)
    g, _ = mlp_ref.gradient(p, x, t, match_device=False)
    eps = 1e-6
    idx = list(rng.choice(5313, 60, replace=False)) + [5312, 5311, 1030, 1100, 5200]
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mlp_ref.gradient(pp, x, t, False)[1] - mlp_ref.gradient(pm, x, t, False)[1]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


def test_reference_features_by_hand():
    # N=8, queens (0,1), (1,3): next row 2; free columns {0,2,4,5,6,7};
    # down-left attacks at row 2: 1+2=3, 3+1=4 -> {3,4}; down-right: 1-2<0, 3-1=2 -> {2}
    f = mlp_ref.features(8, np.array([[1, 3]]))[0] * 16
    assert list(f[:4]) == [2, 6, 2, 1]
    # row 2 free cells: not in {1,3} cols, not {3,4}, not {2} -> {0,5,6,7} = 4
    assert f[4] == 4
