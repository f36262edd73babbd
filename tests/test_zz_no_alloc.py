"""The allocation-free periodic FFT pressure solve (ins2d_poisson_fft_noalloc)
through the C ABI, against the handle-based solve (ins2d_poisson) of the same
grid: two calls in a row on the shared scratch array, the second on a smaller
grid, each within 1e-12 (max norm, relative).

The test is the caller here: it allocates its device arrays and stream with
the HIP runtime the library itself is linked against (found in this
process's memory map), so no second HIP runtime (such as the one bundled
with torch) enters the process."""
import ctypes
import math

import numpy as np
import pytest

from ins2d_amd import _lib, host
from ins2d_amd.cases import load

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[gpu-test]', *a, flush=True)


def _hip_runtime():
    """The libamdhip64 that libins2d_hip.so loaded (not a copy bundled with
    another package)."""
    _lib.load()
    paths = []
    with open('/proc/self/maps') as fh:
        for line in fh:
            p = line.split()[-1]
            if 'libamdhip64.so' in p and p not in paths:
                paths.append(p)
    rocm = [p for p in paths if p.startswith('/opt/rocm')] or paths
    hip = ctypes.CDLL(rocm[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


H2D, D2H = 1, 2


@pytest.mark.parametrize('first,second', [((1024, 512, 1), (256, 128, 0)), ((512, 1024, 0), (64, 32, 1))])
def test_noalloc_fft_solve_matches_handle_solve(first, second):
    lib = _lib.load()
    hip = _hip_runtime()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for Nx, Ny, order4 in (first, second):
            _say('noalloc fft solve', Nx, Ny, 'order4', order4)
            case = load('taylor_green', Nx=Nx, Ny=Ny, Re=500.0, order4=order4, fft_symbol=order4, noise=1e-2, seed=3)
            solver, opts, _, _ = host.setup(case)
            rng = np.random.default_rng(Nx + Ny)
            f = rng.standard_normal(Nx * Ny)
            f -= f.mean()
            dp_ref, _, _ = solver.poisson(f)
            solver.close()
            hx, hy = 2 * math.pi / Nx, 2 * math.pi / Ny
            nb = 8 * Nx * Ny
            fd, dd = ctypes.c_void_p(), ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(fd), nb) == 0
            assert hip.hipMalloc(ctypes.byref(dd), nb) == 0
            try:
                nan = np.full(Nx * Ny, np.nan)
                assert hip.hipMemcpy(fd, f.ctypes.data_as(ctypes.c_void_p), nb, H2D) == 0
                assert hip.hipMemcpy(dd, nan.ctypes.data_as(ctypes.c_void_p), nb, H2D) == 0
                rc = lib.ins2d_poisson_fft_noalloc(Nx, Ny, hx, hy, order4, order4, fd, dd, stream)
                assert rc == 0, lib.ins2d_last_error(None)
                assert hip.hipStreamSynchronize(stream) == 0
                dp = np.empty(Nx * Ny)
                assert hip.hipMemcpy(dp.ctypes.data_as(ctypes.c_void_p), dd, nb, D2H) == 0
            finally:
                hip.hipFree(fd)
                hip.hipFree(dd)
            err = np.abs(dp - dp_ref).max() / np.abs(dp_ref).max()
            _say('noalloc vs handle', Nx, Ny, f'{err:.2e}')
            assert err < 1e-12
    finally:
        hip.hipStreamDestroy(stream)
