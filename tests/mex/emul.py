"""ctypes driver of the MEX gateways built against the MEX API emulation
(tests/mex/Makefile): call a gateway as MATLAB would, with numpy arrays as
MATLAB arrays (column-major, separate real and imaginary parts). Test
infrastructure only."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p


class MexError(RuntimeError):
    pass


class Gateway:
    def __init__(self, name):
        from marie_amd import lib
        lib.load()                      # the HIP runtime and libmarie_vie.so first
        path = os.path.join(HERE, '_build', '%s_emul.so' % name)
        if not os.path.exists(path):
            raise FileNotFoundError('%s not built (make -C tests/mex)' % path)
        so = ctypes.CDLL(path)
        so.emul_new_double.restype = _P
        so.emul_new_double.argtypes = [ctypes.c_int, _P, _P, _P]
        so.emul_new_string.restype = _P
        so.emul_new_string.argtypes = [ctypes.c_char_p]
        so.emul_free.argtypes = [_P]
        so.emul_numel.restype = ctypes.c_int64
        so.emul_numel.argtypes = [_P]
        so.emul_is_complex.argtypes = [_P]
        so.emul_re.restype = _P
        so.emul_re.argtypes = [_P]
        so.emul_im.restype = _P
        so.emul_im.argtypes = [_P]
        so.emul_dim.argtypes = [_P, ctypes.c_int]
        so.emul_ndim.argtypes = [_P]
        so.emul_call.argtypes = [ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_char_p, ctypes.c_int]
        self.so = so

    def _arg(self, v):
        if isinstance(v, str):
            return self.so.emul_new_string(v.encode())
        a = np.asarray(v)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        dims = np.array(a.shape, np.int64)
        re = np.ascontiguousarray(np.real(a).astype(np.float64).ravel(order='F'))
        im = (np.ascontiguousarray(np.imag(a).astype(np.float64).ravel(order='F'))
              if np.iscomplexobj(a) else None)
        return self.so.emul_new_double(a.ndim, dims.ctypes.data, re.ctypes.data,
                                       im.ctypes.data if im is not None else None)

    def _out(self, p):
        n = self.so.emul_numel(p)
        shape = tuple(self.so.emul_dim(p, d) for d in range(max(2, self.so.emul_ndim(p))))
        re = np.ctypeslib.as_array((ctypes.c_double * n).from_address(self.so.emul_re(p))).copy() if n else \
            np.zeros(0)
        if self.so.emul_is_complex(p):
            im = np.ctypeslib.as_array((ctypes.c_double * n).from_address(self.so.emul_im(p))).copy()
            v = re + 1j * im
        else:
            v = re
        return v.reshape(shape, order='F')

    def __call__(self, *args, nargout=1):
        ins = [self._arg(a) for a in args]
        outs = (_P * max(nargout, 1))()
        inarr = (_P * max(len(ins), 1))(*ins)
        err = ctypes.create_string_buffer(1024)
        try:
            st = self.so.emul_call(nargout, outs, len(ins), inarr, err, 1024)
            if st:
                raise MexError(err.value.decode())
            res = [self._out(outs[i]) for i in range(nargout)]
            for i in range(nargout):
                self.so.emul_free(outs[i])
        finally:
            for p in ins:
                self.so.emul_free(p)
        return res[0] if nargout == 1 else res
