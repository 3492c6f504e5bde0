"""ctypes access to the reference's C++ cubatures built under oracle/_ref.

Test infrastructure only. The geometry of each singular integral and the way
the four triangle-pair calls are summed restate singular_st.m:5-20,
singular_ea.m:51-76 / :78-108 and singular_va.m:118-147 / :149-173.
"""
import ctypes
import os

import numpy as np

from .operators import gauss_1d

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref')
_D = ctypes.POINTER(ctypes.c_double)


def available():
    return os.path.exists(os.path.join(_DIR, 'libref_cubature_nop.so'))


def _lib(name):
    return ctypes.CDLL(os.path.join(_DIR, 'libref_%s.so' % name))


def _p(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


class Ref:
    def __init__(self):
        self.cub = _lib('cubature_nop')
        self.kop = _lib('cubature_kop')
        self.kop.ref_cubature_kop.argtypes = [ctypes.c_int] + [_D] * 6 + [ctypes.c_double, _D]
        self.fst = _lib('directfn_st')
        self.fea = _lib('directfn_ea')
        self.fva = _lib('directfn_va')
        self.dst = _lib('demcem_st')
        self.dea = _lib('demcem_ea')
        self.dva = _lib('demcem_va')
        for f in (self.fst.ref_directfn_st,):
            f.argtypes = [_D, _D, _D, ctypes.c_double, ctypes.c_int, _D]
        self.fea.ref_directfn_ea.argtypes = [_D] * 4 + [ctypes.c_double, ctypes.c_int, _D]
        self.fva.ref_directfn_va.argtypes = [_D] * 5 + [ctypes.c_double, ctypes.c_int, _D]
        self.dst.ref_demcem_st.argtypes = [_D] * 3 + [ctypes.c_double, ctypes.c_int, _D, _D, _D]
        self.dea.ref_demcem_ea.argtypes = ([_D] * 4 + [ctypes.c_double, ctypes.c_int, ctypes.c_int]
                                           + [_D] * 5)
        self.dva.ref_demcem_va.argtypes = ([_D] * 5 + [ctypes.c_double] + [ctypes.c_int] * 3
                                           + [_D] * 7)
        self.cub.ref_cubature_nop.argtypes = ([ctypes.c_int] + [_D] * 6
                                              + [ctypes.c_double, ctypes.c_double, _D])
        self._keep = []

    # --- regular cubature (cubature_nop.cpp) ---------------------------------
    def cubature_nop(self, Np, w, u, v, r_m, r_n, R_faces, ko, dx):
        out = np.zeros(72)
        args = [_p(x) for x in (w, u, v, r_m, r_n, np.asarray(R_faces).ravel(order='F'))]
        self.cub.ref_cubature_nop(int(Np), *[a[1] for a in args], float(ko), float(dx),
                                  out.ctypes.data_as(_D))
        return (out[0::2] + 1j * out[1::2]).reshape(6, 6)

    # --- triangle-pair primitives ------------------------------------------
    def cubature_kop(self, Np, w, u, v, r_m, r_n, R_faces, ko):
        out = np.zeros(6)
        args = [_p(x) for x in (w, u, v, r_m, r_n, np.asarray(R_faces).ravel(order='F'))]
        self.kop.ref_cubature_kop(int(Np), *[a[1] for a in args], float(ko), out.ctypes.data_as(_D))
        return out[0::2] + 1j * out[1::2]

    def _call(self, fn, pts, tail):
        out = np.zeros(2)
        ps = [_p(np.asarray(p, float)) for p in pts]
        fn(*[p[1] for p in ps], *tail, out.ctypes.data_as(_D))
        return out[0] + 1j * out[1]

    def st(self, method, r1, r2, r3, ko, Np):
        if method == 'DIRECTFN':
            return self._call(self.fst.ref_directfn_st, (r1, r2, r3), (ko, Np))
        w, z = gauss_1d(Np)
        zz, zp = _p(z); ww, wp = _p(w)
        return self._call(self.dst.ref_demcem_st, (r1, r2, r3), (ko, Np, zp, wp))

    def ea(self, method, r1, r2, r3, r4, ko, Np):
        if method == 'DIRECTFN':
            return self._call(self.fea.ref_directfn_ea, (r1, r2, r3, r4), (ko, Np))
        w, z = gauss_1d(Np)
        (_, wp), (_, zp) = _p(w), _p(z)
        keep = (_p(w), _p(z))
        return self._call(self.dea.ref_demcem_ea, (r1, r2, r3, r4),
                          (ko, Np, Np, keep[0][1], keep[1][1], keep[0][1], keep[1][1]))

    def va(self, method, r1, r2, r3, r4, r5, ko, Np):
        if method == 'DIRECTFN':
            return self._call(self.fva.ref_directfn_va, (r1, r2, r3, r4, r5), (ko, Np))
        w, z = gauss_1d(Np)
        kw, kz = _p(w), _p(z)
        return self._call(self.dva.ref_demcem_va, (r1, r2, r3, r4, r5),
                          (ko, Np, Np, Np, kw[1], kz[1], kw[1], kz[1], kw[1], kz[1]))

    # --- the five square-pair integrals --------------------------------------
    def singular_terms(self, ko, dx, method='DIRECTFN', Np=20):
        d = dx
        r1 = [0, 0, 0]; r2 = [d, 0, 0]; r3 = [d, d, 0]; r4 = [0, d, 0]
        ST = (self.st(method, r1, r2, r3, ko, Np) + self.ea(method, r1, r3, r2, r4, ko, Np)
              + self.ea(method, r1, r3, r4, r2, ko, Np) + self.st(method, r1, r3, r4, ko, Np))
        r5 = [2 * d, 0, 0]; r6 = [2 * d, d, 0]; r5o = [d, 0, d]; r6o = [d, d, d]
        EAc = (self.ea(method, r2, r3, r1, r5, ko, Np) + self.va(method, r3, r1, r2, r5, r6, ko, Np)
               + self.va(method, r3, r4, r1, r2, r5, ko, Np) + self.va(method, r3, r4, r1, r5, r6, ko, Np))
        EAo = (self.ea(method, r2, r3, r1, r5o, ko, Np) + self.va(method, r3, r1, r2, r5o, r6o, ko, Np)
               + self.va(method, r3, r4, r1, r2, r5o, ko, Np)
               + self.va(method, r3, r4, r1, r5o, r6o, ko, Np))
        r5 = [2 * d, d, 0]; r6 = [2 * d, 2 * d, 0]; r7 = [d, 2 * d, 0]
        r5o = [d, d, d]; r6o = [d, 2 * d, d]; r7o = [d, 2 * d, 0]
        VAc = (self.va(method, r3, r1, r2, r5, r6, ko, Np) + self.va(method, r3, r1, r2, r6, r7, ko, Np)
               + self.va(method, r3, r4, r1, r5, r6, ko, Np) + self.va(method, r3, r4, r1, r6, r7, ko, Np))
        VAo = (self.va(method, r3, r1, r2, r5o, r6o, ko, Np)
               + self.va(method, r3, r1, r2, r6o, r7o, ko, Np)
               + self.va(method, r3, r4, r1, r5o, r6o, ko, Np)
               + self.va(method, r3, r4, r1, r6o, r7o, ko, Np))
        return ST, EAc, EAo, VAc, VAo
