"""FP64 numpy restatement of the learned current denoiser (test-only reference
for the HIP kernels of spectral_rs_amd/csrc/denoise.hip; the reference code
has no such network).

Network: x (4 planes) -> z0 = conv3x3(4->32) + b0 -> a0 = GELU(z0) ->
z_l = conv3x3(32->32)(a_{l-1}) + b_l, a_l = GELU(z_l) (l = 1..3) ->
r = conv3x3(32->4)(a3) + b4, output x + r; "same" convolutions, zero
outside the image; GELU in its tanh form (gelu below).  Parameter vector
layout: per layer w_l [cout][cin*9] then b_l [cout]; then scales [4];
w[co][ci*9 + ky*3 + kx].

With emulate=True every convolution operand (input planes, weights, and in
the backward pass the incoming gradient) is rounded to BF16 (round to nearest
even) as the MFMA kernels do, and pre-activations pass through FP32 where the
kernels store them; sums are FP64.  With emulate=False it is plain FP64,
which the finite-difference test pins.
"""
import numpy as np

IN, CH, LAYERS = 4, 32, 5


def cin(l):
    return IN if l == 0 else CH


def cout(l):
    return IN if l == LAYERS - 1 else CH


SHAPES = []
for _l in range(LAYERS):
    SHAPES += [(f"w{_l}", (cout(_l), cin(_l), 3, 3)), (f"b{_l}", (cout(_l),))]
SHAPES.append(("scale", (IN,)))
NTRAIN = int(sum(np.prod(s) for _, s in SHAPES[:-1]))
NPARAMS = NTRAIN + IN


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


GELU_C, GELU_A = np.sqrt(2.0 / np.pi), 0.044715


def gelu(z):
    """GELU, tanh form: z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) / 2."""
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


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


def _ops(emulate):
    if emulate:
        return bf16, lambda z: gelu(np.asarray(z, np.float32).astype(np.float64))
    return (lambda a: np.asarray(a, np.float64)), gelu


def forward(p, x, emulate=True, periodic_y=False):
    """Returns x + r and (pre-activations z_0..z_3, r)."""
    R, act = _ops(emulate)
    q = unpack(p)
    zs = []
    a = np.asarray(x, np.float64)
    for l in range(LAYERS):
        z = conv(R(a), R(q[f"w{l}"]), q[f"b{l}"], periodic_y)
        if l < LAYERS - 1:
            zs.append(z)
            a = act(z)
        else:
            r = z
    return np.asarray(x, np.float64) + r, (zs, r)


def loss_grad(p, x, t, emulate=True):
    """Mean squared error of forward(x) against t and its gradient (NTRAIN)."""
    R, act = _ops(emulate)
    q = unpack(p)
    y, (zs, r) = forward(p, x, emulate)
    n = y.size
    d = y - t
    loss = float((d * d).sum() / n)
    g = {}
    L = LAYERS
    dz = 2.0 * d / n  # gradient of the last layer's output r
    for l in range(L - 1, -1, -1):
        a_in = np.asarray(x, np.float64) if l == 0 else act(zs[l - 1])
        g[f"w{l}"] = wgrad(R(a_in), R(dz))
        g[f"b{l}"] = dz.sum(axis=(0, 2, 3))
        if l > 0:
            dz = conv_t(R(dz), R(q[f"w{l}"])) * gelu_grad(zs[l - 1])
    return loss, pack_grad(g)


def init_params(seed, scale=1.0):
    """He-style random initial parameters with zero biases, a small last
    layer; channel scales set to `scale`."""
    rng = np.random.default_rng(seed)
    q = {}
    for name, shp in SHAPES:
        if name.startswith("w"):
            fan_in = shp[1] * 9
            q[name] = rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)
        else:
            q[name] = np.zeros(shp)
    q[f"w{LAYERS - 1}"] *= 0.1
    q["scale"][:] = scale
    return np.concatenate([q[name].ravel() for name, _ in SHAPES]).astype(np.float32)
