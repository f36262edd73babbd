"""Pins of the CPU oracle of the learned warm-start network (CPU only)."""
import numpy as np

from oracle import unet_oracle as U


def test_bf16_rounding_known_bits():
    # ties go to the even significand; the 8-bit significand keeps 2^-7 steps
    assert U.bf16(1.0) == 1.0
    assert U.bf16(1.0 + 2.0 ** -8) == 1.0
    assert U.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert U.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert U.bf16(-3.0) == -3.0
    assert U.bf16(2.0 ** -130) != 0.0   # subnormals survive as in float32


def test_conv_brute_force():
    rng = np.random.default_rng(1)
    H, W, ci, co = 5, 6, 3, 2
    x = rng.standard_normal((H, W, ci))
    w = rng.standard_normal((9, co, ci))
    b = rng.standard_normal(co)
    out = U.conv(x, w, b, 3)
    ref = np.zeros((H, W, co))
    for y in range(H):
        for xx in range(W):
            for o in range(co):
                s = b[o]
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        yy, xs = y + dy, xx + dx
                        if 0 <= yy < H and 0 <= xs < W:
                            s += x[yy, xs] @ w[3 * (dy + 1) + dx + 1, o]
                ref[y, xx, o] = s
    assert np.allclose(out, ref, rtol=0, atol=1e-12)


def test_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 12, 3))
    w = rng.standard_normal((9, 4, 3))
    g = rng.standard_normal((8, 12, 4))
    lhs = np.sum(U.conv(x, w, None, 3) * g)
    assert np.isclose(lhs, np.sum(x * U.conv_dgrad(g, w, 3)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * U.conv_wgrad(g, x, 3)), rtol=1e-12)
    h = rng.standard_normal((4, 6, 3))
    assert np.isclose(np.sum(U.pool(x) * h), np.sum(x * U.pool_bwd(h)), rtol=1e-12)
    assert np.isclose(np.sum(U.up(h) * x), np.sum(h * U.up_bwd(x)), rtol=1e-12)


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    th = U.init_params(7).astype(np.float64)
    sl, n = U.layer_slices()
    for (_, _, bs) in sl:   # nonzero biases so that every term is exercised
        th[bs] = 0.05 * rng.standard_normal(bs.stop - bs.start)
    fs = [rng.standard_normal((16, 24)) for _ in range(2)]
    ps = [rng.standard_normal((16, 24)) for _ in range(2)]
    L, g = U.loss_and_grad(th, fs, ps, 0.5)
    for k in range(4):
        d = rng.standard_normal(n)
        eps = 1e-7   # small enough that no ReLU input changes sign on this seed
        lp, _ = U.loss_and_grad(th + eps * d, fs, ps, 0.5)
        lm, _ = U.loss_and_grad(th - eps * d, fs, ps, 0.5)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g @ d)
    # single coordinates of the first and last layers
    for idx in (0, sl[0][2].start, sl[14][0].start, n - 1):
        e = np.zeros(n)
        e[idx] = 1.0
        lp, _ = U.loss_and_grad(th + 1e-7 * e, fs, ps, 0.5)
        lm, _ = U.loss_and_grad(th - 1e-7 * e, fs, ps, 0.5)
        assert abs((lp - lm) / 2e-7 - g[idx]) <= 1e-6 * max(1.0, abs(g[idx]))


def test_scale_equivariance():
    """p0(a f) = a p0(f) for a > 0 (input normalised by rms)."""
    rng = np.random.default_rng(4)
    th = U.init_params(3)
    f = rng.standard_normal((16, 16))
    p1 = U.predict(th, f, 2.0)
    p2 = U.predict(th, 7.5 * f, 2.0)
    assert np.allclose(p2, 7.5 * p1, rtol=1e-12, atol=1e-14)


def test_coarse_input_gradient_and_equivariance():
    """With the coarse-grid solution as second input: the gradient against
    central finite differences, p0 = xc + out_scale rms(f) y, and
    p0(a f, a xc) = a p0(f, xc); the input channel matters (a planted zero
    coarse input changes the output)."""
    rng = np.random.default_rng(5)
    th = U.init_params(9).astype(np.float64)
    sl, n = U.layer_slices()
    for (_, _, bs) in sl:
        th[bs] = 0.05 * rng.standard_normal(bs.stop - bs.start)
    fs = [rng.standard_normal((16, 24)) for _ in range(2)]
    ps = [rng.standard_normal((16, 24)) for _ in range(2)]
    xcs = [rng.standard_normal((16, 24)) for _ in range(2)]
    L, g = U.loss_and_grad(th, fs, ps, 0.5, xcs=xcs)
    for k in range(3):
        d = rng.standard_normal(n)
        lp, _ = U.loss_and_grad(th + 1e-7 * d, fs, ps, 0.5, xcs=xcs)
        lm, _ = U.loss_and_grad(th - 1e-7 * d, fs, ps, 0.5, xcs=xcs)
        fd = (lp - lm) / 2e-7
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g @ d)
    f, xc = fs[0], xcs[0]
    y = U.forward(th, f, xc=xc, out_scale=0.5)
    assert np.allclose(U.predict(th, f, 0.5, xc=xc), xc + 0.5 * U.rms(f) * y, rtol=1e-14, atol=1e-14)
    assert np.allclose(U.predict(th, 3.5 * f, 0.5, xc=3.5 * xc), 3.5 * U.predict(th, f, 0.5, xc=xc),
                       rtol=1e-12, atol=1e-13)
    assert np.abs(U.forward(th, f, xc=xc, out_scale=0.5) - U.forward(th, f)).max() > 1e-6
