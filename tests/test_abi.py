"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/ins2d.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from ins2d_amd import _lib


def _declared():
    src = open(os.path.join(ROOT, 'include', 'ins2d.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(ins2d_[a-z_0-9]+)\s*\(', src)))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from ins2d_amd import build
        build.build()
    return ctypes.CDLL(_lib.LIB_PATH)


def test_header_declares_the_bound_entry_points():
    assert _declared() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol(lib):
    for name in _declared():
        assert hasattr(lib, name), name


def test_binding_loads_and_reports_create_errors(lib):
    L = _lib.load()
    d = _lib.CaseDesc()
    d.Nx = 1   # invalid grid: rejected before any device call
    h = ctypes.c_void_p()
    assert L.ins2d_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b'bad grid' in L.ins2d_last_error(None)


def test_noalloc_fft_solve_rejects_oversized_grids(lib):
    """The allocation-free FFT solve checks its sizes before any device call:
    a grid whose half spectrum exceeds the fixed scratch array is refused
    with a clear error (no GPU needed for the refusal)."""
    L = _lib.load()
    dummy = ctypes.c_void_p(8)   # never dereferenced on the refusal path
    rc = L.ins2d_poisson_fft_noalloc(2048, 2048, 1.0, 1.0, 0, 0, dummy, dummy, None)
    assert rc == 1
    assert b'exceeds the scratch capacity' in L.ins2d_last_error(None)
    rc = L.ins2d_poisson_fft_noalloc(48, 64, 1.0, 1.0, 0, 0, dummy, dummy, None)
    assert rc == 4
    assert b'powers of two' in L.ins2d_last_error(None)
