"""Drive matlab/ins2d_mex.c through the MEX stand-in of tests/native/mex_mock.c
(libmex_mock.so, built by ins2d_amd.build): arguments are built the way
MATLAB passes them (column vectors, scalars, strings, cell arrays), a MATLAB
error raised by the gateway comes back as MexError(identifier: message)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'native', 'libmex_mock.so')
vp = ctypes.c_void_p


class MexError(RuntimeError):
    pass


class Cell(list):
    """A MATLAB cell array argument (anything else numeric becomes a double
    column vector, a str a char array)."""


class Mex:
    def __init__(self):
        if not os.path.exists(LIB):
            raise RuntimeError(f'{LIB} is not built: run `python -m ins2d_amd.build`')
        L = ctypes.CDLL(LIB)
        L.mock_double.restype = vp
        L.mock_double.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
        L.mock_string.restype = vp
        L.mock_string.argtypes = [ctypes.c_char_p]
        L.mock_cell.restype = vp
        L.mock_cell.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.mock_data.restype = ctypes.POINTER(ctypes.c_double)
        L.mock_data.argtypes = [vp]
        L.mock_numel.restype = ctypes.c_size_t
        L.mock_numel.argtypes = [vp]
        L.mock_free.argtypes = [vp]
        L.mock_error.restype = ctypes.c_char_p
        L.mock_call.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        self.L = L

    def _arg(self, a, keep):
        if isinstance(a, str):
            p = self.L.mock_string(a.encode())
        elif isinstance(a, Cell):
            items = [self._arg(x, keep) for x in a]
            arr = (vp * max(len(items), 1))(*items)
            p = self.L.mock_cell(arr, len(items))
        else:
            x = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order='F'))
            p = self.L.mock_double(x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), x.size)
        keep.append(p)
        return p

    def __call__(self, *args, nout=1):
        keep = []
        prhs = (vp * len(args))(*[self._arg(a, keep) for a in args])
        plhs = (vp * max(nout, 1))()
        rc = self.L.mock_call(nout, plhs, len(args), prhs)
        for p in keep:
            self.L.mock_free(p)
        if rc:
            raise MexError(self.L.mock_error().decode())
        out = []
        for k in range(nout):
            p = plhs[k]
            if not p:
                out.append(None)
                continue
            n = self.L.mock_numel(p)
            out.append(np.ctypeslib.as_array(self.L.mock_data(p), shape=(n,)).copy())
            self.L.mock_free(p)
        return out[0] if nout == 1 else out
