"""The host assembly of the fourth-order operators on grids with walls
(ins2d_amd/csrc/ops4.cpp) against the oracle, matrix by matrix: the face
averages and interpolations of the eight convective products, the
convective differences, diffusion, divergence, gradient, pressure matrix,
volumes and the boundary vectors at sampled boundary values.

The assembly runs on the CPU through the host-only harness
tests/native/libw4_dump.so (tests/native/w4_dump.cpp, built by
ins2d_amd.build).  Cases: the inviscid cavity of the reference
(case_files/LDC_inviscid), and two made-up cases that reach every boundary
vector -- walls all round on a tanh-stretched grid with nonzero normal and
tangential wall velocities, and walls in x with periodic y."""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp

from ins2d_amd import host
from ins2d_amd.cases import load
from oracle import ins2d_oracle as O
from w4_cases import wall_case

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'native', 'libw4_dump.so')
dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
KIND = {'dir': 0, 'per': 1, 'pres': 2, 'sym': 3}


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f'{LIB} is not built: run `python -m ins2d_amd.build`')
    L = ctypes.CDLL(LIB)
    L.w4h_build.restype = ctypes.c_void_p
    L.w4h_build.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ip, ip, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
    L.w4h_free.argtypes = [ctypes.c_void_p]
    L.w4h_matrix.restype = ctypes.c_long
    L.w4h_matrix.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ip, ip, dp]
    L.w4h_vector.restype = ctypes.c_long
    L.w4h_vector.argtypes = [ctypes.c_void_p, ctypes.c_char_p, dp]
    L.w4h_sizes.argtypes = [ctypes.c_void_p, ip]
    L.w4h_bc.argtypes = [ctypes.c_void_p, ctypes.POINTER(dp), dp, dp, dp, dp]
    return L


def _P(a):
    return a.ctypes.data_as(dp)


CASES = {
    'LDC_inviscid': lambda: load('LDC_inviscid'),
    'walls_stretched': lambda: wall_case(False),
    'walls_x_periodic_y': lambda: wall_case(True),
}


class Assembly:
    def __init__(self, case, opts):
        self.L = L = _lib()
        g, d = opts.grid, opts.discretization
        bct = case.BCtype()
        bu = np.array([KIND[bct['u'][s]] for s in ('left', 'right', 'low', 'up')], np.int32)
        bv = np.array([KIND[bct['v'][s]] for s in ('left', 'right', 'low', 'up')], np.int32)
        x, y = np.ascontiguousarray(g.x), np.ascontiguousarray(g.y)
        err = ctypes.create_string_buffer(256)
        self.h = L.w4h_build(g.Nx, g.Ny, _P(x), _P(y), bu.ctypes.data_as(ip), bv.ctypes.data_as(ip),
                             float(opts.fluid.Re), d.alfa, d.beta, err, 256)
        assert self.h, err.value
        sz = np.zeros(12, np.int32)
        L.w4h_sizes(self.h, sz.ctypes.data_as(ip))
        self.sizes = sz

    def matrix(self, name, shape):
        k = self.L.w4h_matrix(self.h, name.encode(), None, None, None)
        r, c, v = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(k)
        self.L.w4h_matrix(self.h, name.encode(), r.ctypes.data_as(ip), c.ctypes.data_as(ip), _P(v))
        return sp.coo_matrix((v, (r, c)), shape=shape).tocsr()

    def vector(self, name):
        n = self.L.w4h_vector(self.h, name.encode(), None)
        out = np.zeros(n)
        self.L.w4h_vector(self.h, name.encode(), _P(out))
        return out

    def bc(self, samples):
        Nf, NV, Np = int(self.sizes[10]), int(self.sizes[8]), int(self.sizes[9])
        arr = (dp * 12)(*[_P(s) for s in samples])
        out = [np.zeros(Nf), np.zeros(Nf), np.zeros(NV), np.zeros(Np)]
        assert self.L.w4h_bc(self.h, arr, *[_P(o) for o in out]) == 0
        return out

    def close(self):
        self.L.w4h_free(self.h)


@pytest.fixture(scope='module', params=sorted(CASES))
def built(request):
    case = CASES[request.param]()
    opts = O.build(case)
    A = Assembly(case, opts)
    yield request.param, case, opts, A
    A.close()


def _close(name, got, ref, rel=1e-13):
    got, ref = sp.csr_matrix(got), sp.csr_matrix(ref)
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    err = abs(got - ref).max() if got.nnz + ref.nnz else 0.0
    scale = abs(ref).max() if ref.nnz else 1.0
    assert err <= rel * scale, f'{name}: max |diff| {err:.3e} (scale {scale:.3e})'


def test_sizes(built):
    _, _, opts, A = built
    g = opts.grid
    Nx, Ny, nux, nuy, nvx, nvy, Nu, Nv, NV, Np, Nf, singular = (int(v) for v in A.sizes)
    assert (Nx, Ny, nux, nuy, nvx, nvy) == (g.Nx, g.Ny, g.Nux_in, g.Nuy_in, g.Nvx_in, g.Nvy_in)
    assert (Nu, Nv, NV, Np) == (g.Nu, g.Nv, g.NV, g.Np)
    assert singular == 1   # walls and periodic ends only: A has the constant null vector
    np.testing.assert_array_equal(A.vector('xin'), g.xin)
    np.testing.assert_array_equal(A.vector('yin'), g.yin)


def test_operators_match_oracle(built):
    _, _, opts, A = built
    g, d = opts.grid, opts.discretization
    NV, Np, Nf = int(A.sizes[8]), int(A.sizes[9]), int(A.sizes[10])
    a = d.alfa
    Af = sp.bmat([[d.Au_ux, None], [d.Au_uy, None], [None, d.Av_vx], [None, d.Av_vy],
                  [d.Au_ux3, None], [d.Au_uy3, None], [None, d.Av_vx3], [None, d.Av_vy3]])
    If = sp.bmat([[d.Iu_ux, None], [None, d.Iv_uy], [d.Iu_vx, None], [None, d.Iv_vy],
                  [d.Iu_ux3, None], [None, d.Iv_uy3], [d.Iu_vx3, None], [None, d.Iv_vy3]])
    Cw = sp.bmat([[a * d.Cux, a * d.Cuy, None, None, -d.Cux3, -d.Cuy3, None, None],
                  [None, None, a * d.Cvx, a * d.Cvy, None, None, -d.Cvx3, -d.Cvy3]])
    _close('Af', A.matrix('Af', (Nf, NV)), Af)
    _close('If', A.matrix('If', (Nf, NV)), If)
    _close('Cw', A.matrix('Cw', (NV, Nf)), Cw)
    _close('Diff', A.matrix('Diff', (NV, NV)), sp.block_diag([d.Diffu, d.Diffv]))
    _close('M', A.matrix('M', (Np, NV)), d.M)
    _close('G', A.matrix('G', (NV, Np)), d.G)
    _close('A', A.matrix('A', (Np, Np)), d.A)
    np.testing.assert_allclose(A.vector('Om'), g.Om, rtol=1e-15, atol=0)


def test_boundary_vectors_match_oracle(built):
    name, case, opts, A = built
    d = opts.discretization
    yA, yI, yD, yM = A.bc(host.bc_samples(case, opts, 0.0))
    yA_o = np.r_[d.yAu_ux, d.yAu_uy, d.yAv_vx, d.yAv_vy, d.yAu_ux3, d.yAu_uy3, d.yAv_vx3, d.yAv_vy3]
    yI_o = np.r_[d.yIu_ux, d.yIv_uy, d.yIu_vx, d.yIv_vy, d.yIu_ux3, d.yIv_uy3, d.yIu_vx3, d.yIv_vy3]
    for nm, got, ref in (('yA', yA, yA_o), ('yI', yI, yI_o), ('yDiff', yD, np.r_[d.yDiffu, d.yDiffv]),
                         ('yM', yM, d.yM)):
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() <= 1e-13 * scale, (nm, np.abs(got - ref).max(), scale)
    if name != 'LDC_inviscid':
        # the made-up cases reach every boundary vector
        for nm, v in (('yA', yA), ('yI', yI), ('yDiff', yD), ('yM', yM)):
            assert np.abs(v).max() > 1e-3, nm


def test_unsupported_boundaries_rejected():
    L = _lib()
    x = np.linspace(0, 1, 9)
    err = ctypes.create_string_buffer(256)
    for bu, bv in (([2, 0, 0, 0], [0, 0, 0, 0]), ([0, 0, 3, 0], [0, 0, 0, 0]), ([1, 0, 0, 0], [1, 0, 0, 0])):
        bu, bv = np.array(bu, np.int32), np.array(bv, np.int32)
        h = L.w4h_build(8, 8, _P(x), _P(x), bu.ctypes.data_as(ip), bv.ctypes.data_as(ip), 100.0, 81.0, 9 / 8,
                        err, 256)
        assert not h
        assert b'fourth order' in err.value
