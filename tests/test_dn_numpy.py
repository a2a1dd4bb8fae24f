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
    # parameters of every layer, weights and biases (GELU is smooth, so no
    # kink can spoil a central difference)
    o, idx = 0, []
    for name, shp in D.SHAPES[:-1]:
        n = int(np.prod(shp))
        idx += list(o + rng.choice(n, min(n, 8), replace=False))
        o += n
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (D.loss_grad(pp, x, t, False)[0] - D.loss_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, (k, fd, g[k])
    assert len(idx) >= 60


def test_gelu_closed_forms():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    from math import erf, sqrt, tanh
    # the tanh form: z sigmoid(2u), the same numbers by another formula
    u = np.sqrt(2 / np.pi) * (z + 0.044715 * z ** 3)
    assert np.allclose(D.gelu(z), z / (1 + np.exp(-2 * u)), rtol=1e-14, atol=1e-300)
    # within 1e-3 of the exact (erf) GELU on this range (Hendrycks & Gimpel)
    exact = np.array([0.5 * v * (1 + erf(v / sqrt(2))) for v in z])
    assert np.max(np.abs(D.gelu(z) - exact)) < 1e-3
    assert D.gelu(0.0) == 0.0 and D.gelu_grad(0.0) == 0.5
    h = 1e-6
    assert np.allclose(D.gelu_grad(z), (D.gelu(z + h) - D.gelu(z - h)) / (2 * h), rtol=1e-8)
    assert D.NTRAIN == 32 * 36 + 32 + 3 * (32 * 288 + 32) + 4 * 288 + 4
