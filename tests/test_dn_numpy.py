"""The FP64 numpy denoiser (tests/dn_numpy.py) that checks the HIP kernels,
pinned on the CPU: the transposed convolution and the weight gradient are the
adjoints of the convolution, and the analytic gradient equals central finite
differences of the loss."""
import numpy as np

from tests import dn_numpy as D


def test_conv_adjoints():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 7))
    w = rng.standard_normal((4, 3, 3, 3))
    y = rng.standard_normal((2, 4, 5, 7))
    lhs = np.sum(D.conv(x, w) * y)
    assert np.isclose(lhs, np.sum(x * D.conv_t(y, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * D.wgrad(x, y)), rtol=1e-12)


def test_conv_against_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 2, 4, 5))
    w = rng.standard_normal((3, 2, 3, 3))
    b = rng.standard_normal(3)
    out = D.conv(x, w, b)
    o, yy, xx = 2, 0, 4  # an edge pixel: taps outside are zero
    s = b[o]
    for i in range(2):
        for ky in range(3):
            for kx in range(3):
                y, xq = yy + ky - 1, xx + kx - 1
                if 0 <= y < 4 and 0 <= xq < 5:
                    s += w[o, i, ky, kx] * x[0, i, y, xq]
    assert np.isclose(out[0, o, yy, xx], s, rtol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = D.init_params(3).astype(np.float64)
    p[:D.NTRAIN] += 0.05 * rng.standard_normal(D.NTRAIN)  # non-zero biases too
    x = rng.standard_normal((1, 4, 5, 6))
    t = rng.standard_normal((1, 4, 5, 6))
    _, g = D.loss_grad(p, x, t, emulate=False)
    h = 1e-6
    checked = 0
    for k in rng.choice(D.NTRAIN, 60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (D.loss_grad(pp, x, t, False)[0] - D.loss_grad(pm, x, t, False)[0]) / (2 * h)
        # a ReLU kink inside [-h, h] makes finite differences meaningless: skip those
        _, (a1p, a2p, _) = D.forward(pp, x, False)
        _, (a1m, a2m, _) = D.forward(pm, x, False)
        if np.any((a1p > 0) != (a1m > 0)) or np.any((a2p > 0) != (a2m > 0)):
            continue
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, (k, fd, g[k])
        checked += 1
    assert checked >= 50
