"""FP64 numpy restatement of the learned current denoiser (test-only checker
for the HIP kernels of spectral_rs_amd/csrc/denoise.hip; the reference code
has no such network).

Network: x (4 planes) -> conv3x3(4->16)+b, ReLU -> conv3x3(16->16)+b, ReLU ->
conv3x3(16->4)+b = r, output x + r; "same" convolutions, zero outside the
image.  Parameter vector layout: w1 [16][4*9], b1 [16], w2 [16][16*9], b2 [16],
w3 [4][16*9], b3 [4], scales [4]; w[co][ci*9 + ky*3 + kx].

With emulate=True every convolution operand (input planes, weights, and in
the backward pass the incoming gradient) is rounded to BF16 (round to nearest
even) as the MFMA kernels do; sums are FP64.  With emulate=False it is plain
FP64, which the finite-difference test pins.
"""
import numpy as np

NTRAIN, NPARAMS = 3492, 3496
SHAPES = [("w1", (16, 4, 3, 3)), ("b1", (16,)), ("w2", (16, 16, 3, 3)), ("b2", (16,)), ("w3", (4, 16, 3, 3)),
          ("b3", (4,)), ("scale", (4,))]


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack_grad(g):
    return np.concatenate([g[name].ravel() for name, _ in SHAPES[:-1]])


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def conv(x, w, b=None, periodic_y=False):
    """out[b,o,y,x] = b[o] + sum_{i,ky,kx} w[o,i,ky,kx] x[b,i,y+ky-1,x+kx-1], zero outside
    (rows wrap when periodic_y)."""
    B, C, H, W = x.shape
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    if periodic_y:
        xp[:, :, 0, 1:-1] = x[:, :, -1]
        xp[:, :, -1, 1:-1] = x[:, :, 0]
    out = np.zeros((B, w.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oi,biyx->boyx", w[:, :, ky, kx], xp[:, :, ky:ky + H, kx:kx + W])
    if b is not None:
        out += b[None, :, None, None]
    return out


def conv_t(d, w):
    """Transpose of conv: the gradient of conv's input for output gradient d."""
    wt = np.transpose(w[:, :, ::-1, ::-1], (1, 0, 2, 3))
    return conv(d, wt)


def wgrad(x, d):
    """dW[o,i,ky,kx] = sum d[b,o,y,x] x[b,i,y+ky-1,x+kx-1]."""
    B, C, H, W = x.shape
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    g = np.zeros((d.shape[1], C, 3, 3))
    for ky in range(3):
        for kx in range(3):
            g[:, :, ky, kx] = np.einsum("boyx,biyx->oi", d, xp[:, :, ky:ky + H, kx:kx + W])
    return g


def forward(p, x, emulate=True, periodic_y=False):
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    q = unpack(p)
    a1 = np.maximum(conv(R(x), R(q["w1"]), q["b1"], periodic_y), 0)
    a2 = np.maximum(conv(R(a1), R(q["w2"]), q["b2"], periodic_y), 0)
    r = conv(R(a2), R(q["w3"]), q["b3"], periodic_y)
    return np.asarray(x, np.float64) + r, (a1, a2, r)


def loss_grad(p, x, t, emulate=True):
    """Mean squared error of forward(x) against t and its gradient (NTRAIN)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    q = unpack(p)
    y, (a1, a2, r) = forward(p, x, emulate)
    n = y.size
    d = y - t
    loss = float((d * d).sum() / n)
    dr = 2.0 * d / n
    g = {}
    g["w3"] = wgrad(R(a2), R(dr))
    g["b3"] = dr.sum(axis=(0, 2, 3))
    da2 = conv_t(R(dr), R(q["w3"])) * (a2 > 0)
    g["w2"] = wgrad(R(a1), R(da2))
    g["b2"] = da2.sum(axis=(0, 2, 3))
    da1 = conv_t(R(da2), R(q["w2"])) * (a1 > 0)
    g["w1"] = wgrad(R(x), R(da1))
    g["b1"] = da1.sum(axis=(0, 2, 3))
    return loss, pack_grad(g)


def init_params(seed, scale=1.0):
    """He-style random initial parameters with zero biases; channel scales set to `scale`."""
    rng = np.random.default_rng(seed)
    q = {}
    for name, shp in SHAPES:
        if name.startswith("w"):
            fan_in = shp[1] * 9
            q[name] = rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)
        else:
            q[name] = np.zeros(shp)
    q["w3"] *= 0.1
    q["scale"][:] = scale
    return np.concatenate([q[name].ravel() for name, _ in SHAPES]).astype(np.float32)
