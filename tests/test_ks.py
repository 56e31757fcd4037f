"""KS aggregation (host function of the C ABI, no GPU use) against closed
forms and its own central differences."""
import numpy as np

import shellgpu as sg


def test_ks_closed_forms():
    ks, d = sg.ks_aggregate(np.full(3, 2.5), 100.0)
    assert abs(ks - (2.5 + np.log(3.0) / 100.0)) < 1e-15
    assert np.allclose(d, 1.0 / 3.0, rtol=0, atol=1e-16)
    ks, d = sg.ks_aggregate(np.array([0.0, -1000.0]), 100.0)
    assert ks == 0.0 and d[0] == 1.0 and d[1] == 0.0
    # max-like: between max and max + log(n)/rho
    f = np.random.default_rng(1).standard_normal(7)
    ks, d = sg.ks_aggregate(f, 10.0)
    assert f.max() <= ks <= f.max() + np.log(7) / 10.0
    assert abs(d.sum() - 1.0) < 1e-15 and (d > 0).all()


def test_ks_gradient_matches_differences():
    f = np.random.default_rng(2).standard_normal(5) * 0.05
    ks, d = sg.ks_aggregate(f, 100.0)
    h = 1e-6
    for i in range(5):
        e = np.zeros(5)
        e[i] = h
        fd = (sg.ks_aggregate(f + e, 100.0)[0] - sg.ks_aggregate(f - e, 100.0)[0]) / (2 * h)
        assert abs(fd - d[i]) < 1e-7
