"""The MATLAB MEX gateway (matlab/ins2d_mex.c) without a GPU: argument and
handle checks and the error mapping, driven through the MEX stand-in."""
import pytest

from mex_harness import Mex, MexError


@pytest.fixture(scope='module')
def mex():
    return Mex()


def test_gateway_rejects_bad_calls(mex):
    with pytest.raises(MexError, match='INS2D:args: command expected'):
        mex(1.0)
    with pytest.raises(MexError, match='INS2D:args: unknown command'):
        mex('frobnicate', 0)
    with pytest.raises(MexError, match='INS2D:handle: invalid handle'):
        mex('step', 3, 0.0, 0.01, 1)
    with pytest.raises(MexError, match='INS2D:args: create'):
        mex('create', [0, 1], [0, 1])


def test_gateway_maps_library_errors(mex):
    # a one-cell grid is refused by ins2d_create before any device call; the
    # gateway raises it as INS2D:error with the library's message
    with pytest.raises(MexError, match='INS2D:error: .*bad grid'):
        mex('create', [0.0, 1.0], [0.0, 1.0], [0, 0, 0, 0], [0, 0, 0, 0], 100.0, 0,
            [[0.0]], [1.0], [0.0], 1, 1e-8, 100, 1)


def test_gateway_checks_array_kinds(mex):
    from mex_harness import Cell
    with pytest.raises(MexError, match='INS2D:args: x: a double array expected'):
        mex('create', Cell([0.0, 1.0]), [0.0, 1.0], [0, 0, 0, 0], [0, 0, 0, 0], 100.0, 0,
            [[0.0]], [1.0], [0.0], 1, 1e-8, 100, 1)
