"""Host side of the learned Poisson warm start (no reference counterpart).

``WarmStartNet`` wraps the C-ABI network object of libins2d_hip: a U-shaped
CNN evaluated and trained by the library's own HIP kernels (BF16 operands on
the matrix cores, FP32 accumulation).  Device arrays handed to ``train_step``
and ``forward_dev`` are raw device addresses (for example ``tensor.data_ptr()``
of a CUDA tensor); the calls queue work on the null stream.
"""
import ctypes

import numpy as np

from . import _lib
from .host import Ins2dError


class WarmStartNet:
    def __init__(self, H, W):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.ins2d_nn_create(int(H), int(W), ctypes.byref(h))
        if rc != 0:
            raise Ins2dError(rc, self._lib.ins2d_nn_last_error(None).decode())
        self._h = h
        self.H, self.W = int(H), int(W)
        self.nparam = int(self._lib.ins2d_nn_num_params(h))
        self.out_scale = 1.0

    def _check(self, rc):
        if rc != 0:
            raise Ins2dError(rc, self._lib.ins2d_nn_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.ins2d_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta, out_scale):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.size != self.nparam:
            raise ValueError(f'expected {self.nparam} parameters, got {th.size}')
        self.out_scale = float(out_scale)
        self._check(self._lib.ins2d_nn_set_params(self._h, th.ctypes.data_as(_lib.P_float), self.out_scale))

    def init(self, seed=0, out_scale=1.0):
        """He-normal weights, zero biases, fresh Adam state."""
        self.out_scale = float(out_scale)
        self._check(self._lib.ins2d_nn_init(self._h, ctypes.c_ulonglong(int(seed)), self.out_scale))

    def save(self, path):
        np.savez(path, theta=self.get_params(), out_scale=self.out_scale, H=self.H, W=self.W)

    def load(self, path):
        d = np.load(path)
        if int(d['H']) != self.H or int(d['W']) != self.W:
            raise ValueError('saved network is for another grid')
        self.set_params(d['theta'], float(d['out_scale']))

    def get_params(self):
        th = np.empty(self.nparam, dtype=np.float32)
        self._check(self._lib.ins2d_nn_get_params(self._h, th.ctypes.data_as(_lib.P_float)))
        return th

    def get_grad(self):
        g = np.empty(self.nparam, dtype=np.float32)
        self._check(self._lib.ins2d_nn_get_grad(self._h, g.ctypes.data_as(_lib.P_float)))
        return g

    def forward(self, f, xc=None, unfused=False):
        """p0 for a host right-hand side f of shape (H, W) (x fastest) and an
        optional coarse-grid solution xc of the same shape (``unfused``: through
        the training forward)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != (self.H, self.W):
            raise ValueError(f'expected shape {(self.H, self.W)}, got {f.shape}')
        x = None if xc is None else np.ascontiguousarray(xc, dtype=np.float64).reshape(f.shape)
        p = np.empty_like(f)
        fn = self._lib.ins2d_nn_forward_unfused if unfused else self._lib.ins2d_nn_forward
        self._check(fn(self._h, f.ctypes.data_as(_lib.P_double),
                                               None if x is None else x.ctypes.data_as(_lib.P_double),
                                               p.ctypes.data_as(_lib.P_double)))
        return p

    def forward_dev(self, f_ptr, p_ptr, xc_ptr=None):
        self._check(self._lib.ins2d_nn_forward_dev(self._h, ctypes.c_void_p(f_ptr),
                                                   None if xc_ptr is None else ctypes.c_void_p(xc_ptr),
                                                   ctypes.c_void_p(p_ptr), None))

    def train_step(self, f_ptrs, p_ptrs, lr=1e-3, update=True, xc_ptrs=None):
        """Loss over a batch of device (f, p[, xc]) pairs; with ``update`` one
        Adam step.  Returns the mean loss."""
        n = len(f_ptrs)
        fa = (ctypes.c_void_p * n)(*[ctypes.c_void_p(x) for x in f_ptrs])
        pa = (ctypes.c_void_p * n)(*[ctypes.c_void_p(x) for x in p_ptrs])
        xa = None if xc_ptrs is None else (ctypes.c_void_p * n)(*[ctypes.c_void_p(x) for x in xc_ptrs])
        loss = ctypes.c_double()
        self._check(self._lib.ins2d_nn_train_step(self._h, fa, pa, xa, n, ctypes.c_float(lr), 1 if update else 0,
                                                  ctypes.byref(loss), None))
        return loss.value


    def data_reserve(self, count):
        self._check(self._lib.ins2d_nn_data_reserve(self._h, int(count)))

    def data_set(self, k, f, p, xc=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        x = None if xc is None else np.ascontiguousarray(xc, dtype=np.float64)
        if f.size != self.H * self.W or p.size != self.H * self.W or (x is not None and x.size != f.size):
            raise ValueError('training pair has the wrong size')
        self._check(self._lib.ins2d_nn_data_set(self._h, int(k), f.ctypes.data_as(_lib.P_double),
                                                p.ctypes.data_as(_lib.P_double),
                                                None if x is None else x.ctypes.data_as(_lib.P_double)))

    def data_ptr(self, k, which):
        """Device address of pair k's array (0 f, 1 p, 2 xc)."""
        ptr = self._lib.ins2d_nn_data_ptr(self._h, int(k), int(which))
        if not ptr:
            raise IndexError(f'no training pair {k}')
        return int(ptr)

    def data_get(self, k):
        """Host copies (f, p, xc) of pair k."""
        f, p, xc = (np.empty(self.H * self.W) for _ in range(3))
        P = lambda a: a.ctypes.data_as(_lib.P_double)
        self._check(self._lib.ins2d_nn_data_get(self._h, int(k), P(f), P(p), P(xc)))
        return f, p, xc

    def data_combine(self, dst, i, j, a, b):
        """pair dst = a pair_i / rms(f_i) + b pair_j / rms(f_j) (device)."""
        self._check(self._lib.ins2d_nn_data_combine(self._h, int(dst), int(i), int(j), ctypes.c_double(a),
                                                    ctypes.c_double(b)))

    def train(self, idx, lr=1e-3, update=True, relative=False):
        """Loss over the stored pairs ``idx``; with ``update`` one Adam step;
        ``relative``: per-pair loss sum (y - t)^2 / sum t^2."""
        ia = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ctypes.c_double()
        mode = (1 if update else 0) | (2 if relative else 0)
        self._check(self._lib.ins2d_nn_train(self._h, ia.ctypes.data_as(_lib.P_int), ia.size, ctypes.c_float(lr),
                                             mode, ctypes.byref(loss)))
        return loss.value


def layer_check(xa, xb, w, bias, g, ks, relu):
    """Run one convolution layer's forward, input-gradient, weight-gradient
    and bias-gradient kernels on host arrays; returns (out, din, dw, db)."""
    lib = _lib.load()
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    xa, w, bias, g = f32(xa), f32(w), f32(bias), f32(g)
    H, W, ca = xa.shape
    cb = 0 if xb is None else xb.shape[2]
    xb = f32(np.zeros(1) if xb is None else xb)
    cout = w.shape[1]
    out = np.empty((H, W, cout), np.float32)
    din = np.empty((H, W, ca + cb), np.float32)
    dw = np.empty_like(w)
    db = np.empty(cout, np.float32)
    P = lambda a: a.ctypes.data_as(_lib.P_float)
    rc = lib.ins2d_nn_layer_check(H, W, ca, cb, cout, ks, relu, P(xa), P(xb), P(w), P(bias), P(g), P(out), P(din),
                                  P(dw), P(db))
    if rc != 0:
        raise Ins2dError(rc, lib.ins2d_nn_last_error(None).decode())
    return out, din, dw, db


def mfma_probe(A, B):
    """D = A B through one BF16 16x16x16 matrix-core instruction."""
    lib = _lib.load()
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    D = np.empty((16, 16), dtype=np.float32)
    rc = lib.ins2d_nn_mfma_probe(A.ctypes.data_as(_lib.P_float), B.ctypes.data_as(_lib.P_float),
                                 D.ctypes.data_as(_lib.P_float))
    if rc != 0:
        raise Ins2dError(rc, 'mfma probe failed')
    return D
