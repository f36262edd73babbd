"""CPU oracle of the learned Poisson warm-start network (test infrastructure).

This file is a checker: only tests, the smoke check and the benchmark's CPU
leg may import it.  It has no counterpart in the reference (INS2D has no
learned component); it restates, in numpy FP64, the network that the HIP
library evaluates:

  4-level U-shaped CNN, 3x3 convolutions with zero padding, ReLU, 2x2 average
  pooling, nearest x2 upsampling, skip connections by channel concatenation
  (upsampled decoder features first), a final 1x1 linear convolution;
  channels 16/32/64/64; input [f / rms(f), xc / (out_scale rms(f)), 0, 0]
  with xc an optional coarse-grid solution (zero when absent); output y with
  p0 = xc + out_scale * rms(f) * y; loss
  mean((y - (p - xc) / (out_scale rms(f)))^2).

Tensors are (H, W, C) arrays, W fastest over pixels as in the solver's
pressure layout.  Parameters are one flat vector: for each layer the weights
w[t][co][ci] (t = 3 (dy + 1) + (dx + 1)) then the bias.

``bf16=True`` rounds to bfloat16 (round to nearest even) exactly where the
GPU kernels store or load bfloat16: weights, stored activations, pooled and
upsampled tensors, the network input, and every matrix-core operand
(including the FP32 gradients fed to the backward convolutions).  Sums are
FP64 here and FP32 on the GPU.

Parity pins (tests/test_unet_oracle.py): the convolution against a brute-force
loop; pooling/upsampling as mutual adjoints; the analytic gradient of
``loss_and_grad`` against central finite differences of ``forward`` in FP64;
the bf16 rounding against known bit patterns.
"""
import numpy as np

# (level, cin_a, cin_b, cout, ks, relu)
LAYERS = [
    (0, 4, 0, 16, 3, 1), (0, 16, 0, 16, 3, 1),
    (1, 16, 0, 32, 3, 1), (1, 32, 0, 32, 3, 1),
    (2, 32, 0, 64, 3, 1), (2, 64, 0, 64, 3, 1),
    (3, 64, 0, 64, 3, 1), (3, 64, 0, 64, 3, 1),
    (2, 64, 64, 64, 3, 1), (2, 64, 0, 64, 3, 1),
    (1, 64, 32, 32, 3, 1), (1, 32, 0, 32, 3, 1),
    (0, 32, 16, 16, 3, 1), (0, 16, 0, 16, 3, 1),
    (0, 16, 0, 1, 1, 0),
]


def layer_slices():
    """(weight slice, weight shape, bias slice) per layer in the flat vector."""
    out, off = [], 0
    for (_, ca, cb, co, ks, _) in LAYERS:
        nw = ks * ks * co * (ca + cb)
        out.append((slice(off, off + nw), (ks * ks, co, ca + cb), slice(off + nw, off + nw + co)))
        off += nw + co
    return out, off


def num_params():
    return layer_slices()[1]


def init_params(seed=0):
    """He-normal weights, zero biases (float32, as the trainer starts)."""
    rng = np.random.default_rng(seed)
    sl, n = layer_slices()
    th = np.zeros(n, dtype=np.float64)
    for (ws, shp, _), (_, ca, cb, co, ks, relu) in zip(sl, LAYERS):
        fan_in = ks * ks * (ca + cb)
        std = np.sqrt((2.0 if relu else 1.0) / fan_in)
        th[ws] = rng.standard_normal(int(np.prod(shp))) * std
    return th.astype(np.float32)


def bf16(x):
    """Round to bfloat16 (nearest even) via float32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def _shift(x, dy, dx):
    """s[y, x] = x[y + dy, x + dx] with zeros outside."""
    H, W = x.shape[:2]
    s = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    s[yd, xd] = x[ys, xs]
    return s


def _offsets(ks):
    r = ks // 2
    return [(t // ks - r, t % ks - r) for t in range(ks * ks)]


def conv(x, w, b, ks):
    """out[y, x, co] = b[co] + sum_t sum_ci x[y+dy_t, x+dx_t, ci] w[t, co, ci]."""
    out = np.zeros(x.shape[:2] + (w.shape[1],))
    for t, (dy, dx) in enumerate(_offsets(ks)):
        out += _shift(x, dy, dx) @ w[t].T
    if b is not None:
        out += b
    return out


def conv_dgrad(g, w, ks):
    """Adjoint of ``conv`` in its input: din[q] = sum_t g[q - off_t] w[t]^T."""
    din = np.zeros(g.shape[:2] + (w.shape[2],))
    for t, (dy, dx) in enumerate(_offsets(ks)):
        din += _shift(g @ w[t], -dy, -dx)
    return din


def conv_wgrad(g, x, ks):
    """dw[t, co, ci] = sum_p g[p, co] x[p + off_t, ci]."""
    C = g.shape[2]
    dw = np.zeros((ks * ks, C, x.shape[2]))
    g2 = g.reshape(-1, C)
    for t, (dy, dx) in enumerate(_offsets(ks)):
        dw[t] = g2.T @ _shift(x, dy, dx).reshape(-1, x.shape[2])
    return dw


def pool(x):
    H, W, C = x.shape
    return 0.25 * x.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=0), 2, axis=1)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_bwd(g):
    H, W, C = g.shape
    return g.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def _layer_params(theta, R):
    sl, _ = layer_slices()
    th = np.asarray(theta, dtype=np.float64)
    return [(R(th[ws].reshape(shp)), th[bs]) for (ws, shp, bs) in sl]


def rms(f):
    return float(np.sqrt(np.mean(np.asarray(f, dtype=np.float64) ** 2)))


def forward(theta, f, bf16_emulation=False, keep=False, xc=None, out_scale=1.0):
    """Network output y (H, W) for a right-hand side f (H, W) and an optional
    coarse-grid solution xc (H, W)."""
    R = bf16 if bf16_emulation else _ident
    P = _layer_params(theta, R)
    s = max(rms(f), 1e-300)
    H, W = f.shape
    x0 = np.zeros((H, W, 4))
    x0[..., 0] = R(np.float32(np.asarray(f, dtype=np.float64) / s)) if bf16_emulation else f / s
    if xc is not None:
        c = np.asarray(xc, dtype=np.float64).reshape(H, W) / (out_scale * s)
        x0[..., 1] = R(np.float32(c)) if bf16_emulation else c
    T = {'x0': x0, 's': s}

    def layer(l, a, b=None):
        inp = a if b is None else np.concatenate([a, b], axis=2)
        w, bias = P[l]
        ks, relu = LAYERS[l][4], LAYERS[l][5]
        z = conv(R(inp), w, bias, ks)
        out = np.maximum(z, 0.0) if relu else z
        if l < 14:
            out = R(out)
        T['in%d' % l] = inp
        T['a%d' % l] = out
        return out

    a0 = layer(0, x0)
    a1 = layer(1, a0)
    T['pool1'] = p1 = R(pool(a1))
    a2 = layer(2, p1)
    a3 = layer(3, a2)
    T['pool2'] = p2 = R(pool(a3))
    a4 = layer(4, p2)
    a5 = layer(5, a4)
    T['pool3'] = p3 = R(pool(a5))
    a6 = layer(6, p3)
    a7 = layer(7, a6)
    a8 = layer(8, up(a7), a5)
    a9 = layer(9, a8)
    a10 = layer(10, up(a9), a3)
    a11 = layer(11, a10)
    a12 = layer(12, up(a11), a1)
    a13 = layer(13, a12)
    y = layer(14, a13)[..., 0]
    return (y, T) if keep else y


def predict(theta, f, out_scale, bf16_emulation=False, xc=None):
    """p0 = xc + out_scale * rms(f) * y(f, xc) (xc = 0 when absent)."""
    y = forward(theta, f, bf16_emulation, xc=xc, out_scale=out_scale)
    base = 0.0 if xc is None else np.asarray(xc, dtype=np.float64).reshape(y.shape)
    return base + out_scale * rms(f) * y


def loss_and_grad(theta, fs, ps, out_scale, bf16_emulation=False, xcs=None):
    """Mean over the batch of mean((y - (p - xc) / (out_scale rms f))^2) and
    its gradient with respect to the flat parameter vector."""
    R = bf16 if bf16_emulation else _ident
    P = _layer_params(theta, R)
    sl, n = layer_slices()
    grad = np.zeros(n)
    total = 0.0
    for k, (f, p) in enumerate(zip(fs, ps)):
        xc = None if xcs is None else xcs[k]
        y, T = forward(theta, f, bf16_emulation, keep=True, xc=xc, out_scale=out_scale)
        npx = y.size
        base = 0.0 if xc is None else np.asarray(xc, dtype=np.float64).reshape(y.shape)
        tt = (np.asarray(p, dtype=np.float64) - base) / (out_scale * T['s'])
        d = y - tt
        total += float(np.sum(d * d) / npx)
        g_out = (2.0 * d / npx)[..., None]

        def back(l, g):
            """g: gradient at the layer output (pre-activation already masked)."""
            ws, shp, bs = sl[l]
            ks = LAYERS[l][4]
            inp = T['in%d' % l]
            grad[ws] += conv_wgrad(R(g), R(inp), ks).ravel()
            grad[bs] += g.reshape(-1, g.shape[2]).sum(axis=0)
            return conv_dgrad(R(g), P[l][0], ks)

        def relu_mask(g, l):
            return g * (T['a%d' % l] > 0)

        def split(g, ca):
            return g[..., :ca], g[..., ca:]

        g = back(14, g_out)
        g = back(13, relu_mask(g, 13))
        g = back(12, relu_mask(g, 12))
        d_up1, d_s1 = split(g, 32)
        g = up_bwd(d_up1)
        g = back(11, relu_mask(g, 11))
        g = back(10, relu_mask(g, 10))
        d_up2, d_s2 = split(g, 64)
        g = up_bwd(d_up2)
        g = back(9, relu_mask(g, 9))
        g = back(8, relu_mask(g, 8))
        d_up3, d_s3 = split(g, 64)
        g = up_bwd(d_up3)
        g = back(7, relu_mask(g, 7))
        g = back(6, relu_mask(g, 6))
        d_s3 = d_s3 + pool_bwd(g)
        g = back(5, relu_mask(d_s3, 5))
        g = back(4, relu_mask(g, 4))
        d_s2 = d_s2 + pool_bwd(g)
        g = back(3, relu_mask(d_s2, 3))
        g = back(2, relu_mask(g, 2))
        d_s1 = d_s1 + pool_bwd(g)
        g = back(1, relu_mask(d_s1, 1))
        back(0, relu_mask(g, 0))
    nb = len(fs)
    return total / nb, grad / nb
