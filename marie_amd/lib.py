"""ctypes binding of libmarie_vie.so (the C-ABI declared in include/marie_vie.h).

The library is the only compute path: loading fails loudly when the shared
object is missing or no GPU is present; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmarie_vie.so')

PREC_F64 = 1
PREC_F32 = 2
OP_AN = 0
OP_N = 1
OP_L = 2
OP_K = 3
OP_P = 4            # the JVIE preconditioner: mv_AN with (1./Mr, 1./Mr - 1)
PRECOND_NONE = 0
PRECOND_RIGHT = 1
PRECOND_LEFT = 2

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ('marie_vie_version', _I, []),
    ('marie_vie_last_error', ctypes.c_char_p, []),
    ('marie_vie_create', _I, [ctypes.POINTER(_P), _I, _I, _I, _I, _I, _I, _I, _I]),
    ('marie_vie_destroy', _I, [_P]),
    ('marie_vie_set_stream', _I, [_P, _P]),
    ('marie_vie_set_fN', _I, [_P, _P]),
    ('marie_vie_set_fN_device', _I, [_P, _P]),
    ('marie_vie_get_fN', _I, [_P, _P]),
    ('marie_vie_assemble_N', _I, [_P, _D, _D, _P]),
    ('marie_vie_singular_terms', _I, [_D, _D, _P]),
    ('marie_vie_assemble_type', _I, [_P, _I, _D, _D, _P, _I, _P]),
    ('marie_vie_singular_terms_directfn', _I, [_D, _D, _I, _P]),
    ('marie_vie_set_fK', _I, [_P, _P]),
    ('marie_vie_get_fK', _I, [_P, _P]),
    ('marie_vie_assemble_K', _I, [_P, _D, _D]),
    ('marie_vie_set_material', _I, [_P, _P, _P, _D]),
    ('marie_vie_set_index', _I, [_P, _P, _I64, _I]),
    ('marie_vie_apply', _I, [_P, _I, _I, _P, _P, _I]),
    ('marie_vie_apply_device', _I, [_P, _I, _I, _I, _P, _P, _I]),
    ('marie_vie_apply_device_noalloc', _I, [_P, _I, _I, _I, _P, _P, _I]),
    ('marie_vie_noalloc_scratch_bytes', ctypes.c_size_t, []),
    ('marie_vie_gmres_device', _I, [_P, _P, _P, _I, _D, _I, _I, _D, ctypes.POINTER(_I), _DP,
                                    ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres', _I, [_P, _P, _P, _I, _D, _I, _I, _D, ctypes.POINTER(_I), _DP,
                             ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres_adjoint_device', _I, [_P, _P, _P, _I, _D, _I, _I, _D, ctypes.POINTER(_I), _DP,
                                            ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres_precond_device', _I, [_P, _P, _P, _I, _D, _I, _I, ctypes.POINTER(_I), _DP,
                                            ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres_precond', _I, [_P, _P, _P, _I, _D, _I, _I, ctypes.POINTER(_I), _DP,
                                     ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres_dr_device', _I, [_P, _P, _P, _I, _D, _I, _I, _I, ctypes.POINTER(_I), _DP,
                                       ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_gmres_dr', _I, [_P, _P, _P, _I, _D, _I, _I, _I, ctypes.POINTER(_I), _DP,
                                ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_bicgstab_device', _I, [_P, _P, _P, _D, _I, _I, ctypes.POINTER(_I), _DP, _DP, _DP, _I,
                                       ctypes.POINTER(_I)]),
    ('marie_vie_bicgstab', _I, [_P, _P, _P, _D, _I, _I, ctypes.POINTER(_I), _DP, _DP, _DP, _I,
                                ctypes.POINTER(_I)]),
    ('marie_vie_material_gradient_device', _I, [_P, _P, _P, _P, _D, _P, _P]),
    ('marie_vie_block_gmres_device', _I, [_P, _P, _P, _I, _I, _D, _I, _I, _D, _I, ctypes.POINTER(_I), _DP,
                                          ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_block_gmres', _I, [_P, _P, _P, _I, _I, _D, _I, _I, _D, _I, ctypes.POINTER(_I), _DP,
                                   ctypes.POINTER(_I), _DP, _I, ctypes.POINTER(_I)]),
    ('marie_vie_block_gmres_batch', _I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                         ctypes.POINTER(_I), _I, _D, _I, _I, _D, _I, ctypes.POINTER(_I), _DP,
                                         ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    ('marie_vie_dgf_fields', _I, [_P, _P, _P, _I, _P, _I64, _D, _D, _P, _P, _I]),
    ('marie_vie_dgf_fields_device', _I, [_P, _P, _P, _I, _P, _I64, _D, _D, _P, _P, _P]),
    ('marie_vie_autotune', _I, [_P, _I, _I, ctypes.POINTER(_I), _DP]),
    ('marie_vie_set_tiles', _I, [_P, _I, ctypes.POINTER(_I)]),
    ('marie_vie_enable_timing', _I, [_P, _I]),
    ('marie_vie_get_timing', _I, [_P, _DP, ctypes.POINTER(_I64)]),
]

_lib = None


class MarieError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MarieError('libmarie_vie.so is not built: run `python -m marie_amd.build` '
                         '(or __graft_entry__.build())')
    # A process must hold one HIP runtime. PyTorch ships its own copy
    # (soname libamdhip64.so.7, which then also serves this library); loaded
    # after this library, torch would bring up a second runtime that sees no
    # GPU. So torch, when installed, is imported first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = _lib.marie_vie_last_error().decode(errors='replace')
        raise MarieError('marie_vie error %d: %s' % (status, msg))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _cplx(a):
    """C-contiguous complex128 view in MATLAB column-major element order."""
    a = np.asarray(a, dtype=np.complex128)
    return np.ascontiguousarray(a.ravel(order='F'))


def _vec_arg(b, x0, nidx, name='b'):
    """b and x0 (or zeros) as contiguous complex128 vectors of exactly nidx
    entries: the C entry points read and write nidx values through them."""
    b = np.ascontiguousarray(np.asarray(b, np.complex128).ravel())
    if b.size != nidx:
        raise ValueError('%s must hold nidx = %d entries (got %d)' % (name, nidx, b.size))
    if x0 is None:
        return b, np.zeros_like(b)
    x = np.ascontiguousarray(np.asarray(x0, np.complex128).ravel()).copy()
    if x.size != nidx:
        raise ValueError('x0 must hold nidx = %d entries (got %d)' % (nidx, x.size))
    return b, x


def dgf_fields(seg_center, seg_J, port_start, points, ko, omega_mu, device=0):
    """E, H (nport x npoints x 3 complex) of current elements (eval_DGF.m) on the GPU.

    seg_center: nseg x 3 midpoints; seg_J: nseg x 3 complex current elements;
    port_start: nport+1 segment offsets; points: npoints x 3.
    """
    lib = load()
    sc = np.ascontiguousarray(np.asarray(seg_center, np.float64).reshape(-1, 3))
    sj = np.ascontiguousarray(np.asarray(seg_J, np.complex128).reshape(-1, 3))
    ps = np.ascontiguousarray(np.asarray(port_start, np.int32))
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    npts = pts.shape[0]
    cm = np.ascontiguousarray(pts.T)                 # column-major npoints x 3
    nport = ps.size - 1
    E = np.empty(nport * 3 * npts, np.complex128)
    H = np.empty_like(E)
    check(lib.marie_vie_dgf_fields(_ptr(sc), _ptr(sj), _ptr(ps), nport, _ptr(cm), npts, float(ko), float(omega_mu),
                                   _ptr(E), _ptr(H), device))
    return (E.reshape(nport, 3, npts).transpose(0, 2, 1), H.reshape(nport, 3, npts).transpose(0, 2, 1))


BLOCK_GRAPH = 1
BLOCK_CGS2 = 2
BLOCK_COLUMNS = 4
BLOCK_DEVICE = 8


def _block_options(use_graph, cgs_passes, columns=False, device_steps=False):
    if cgs_passes not in (1, 2):
        raise ValueError('cgs_passes must be 1 or 2')
    return ((BLOCK_GRAPH if use_graph else 0) | (BLOCK_CGS2 if cgs_passes == 2 else 0) |
            (BLOCK_COLUMNS if columns else 0) | (BLOCK_DEVICE if device_steps else 0))


def block_gmres_batch(ops, B_ptrs, X_ptrs, nrhs, restart=20, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4,
                      use_graph=False, cgs_passes=1, columns=False, device_steps=False):
    """Block GMRES of several problems at once (one handle each, device
    buffers B_ptrs[i], X_ptrs[i] of nrhs[i] stacked columns); columns=True
    solves each problem's columns one at a time with GMRES in `mode` (0-3)
    instead; device_steps=True runs every block step on the device and drives
    all problems from one host thread (with use_graph: each step a HIP graph).
    Returns (flags per problem, relres per problem, [steps, cycles] per
    problem)."""
    n = len(ops)
    tot = int(sum(nrhs))
    P = ctypes.c_void_p
    opa = (P * n)(*[o._h.value for o in ops])
    ba = (P * n)(*B_ptrs)
    xa = (P * n)(*X_ptrs)
    na = (ctypes.c_int * n)(*nrhs)
    flag = (ctypes.c_int * tot)()
    relres = (ctypes.c_double * tot)()
    it = (ctypes.c_int * (2 * n))()
    status = (ctypes.c_int * n)()
    check(load().marie_vie_block_gmres_batch(opa, n, ba, xa, na, restart, tol, maxit, mode, inner_tol,
                                             _block_options(use_graph, cgs_passes, columns, device_steps), flag,
                                             relres, it, status))
    fl, rr, its, o = [], [], [], 0
    for i in range(n):
        fl.append(np.array(flag[o:o + nrhs[i]]))
        rr.append(np.array(relres[o:o + nrhs[i]]))
        its.append([it[2 * i], it[2 * i + 1]])
        o += nrhs[i]
    return fl, rr, its


def _largest_prime_factor(n):
    f, p = 1, 2
    while p * p <= n:
        while n % p == 0:
            f, n = p, n // p
        p += 1
    return max(f, n) if n > 1 else f


def circulant_side(n):
    """Circulant side for a grid side of n voxels: 2n, as getOPERATORS.m
    builds it, when 2n factors into primes up to 13 (the FFT radices); else
    the smallest even side >= 2n with factors up to 7. Any side >= 2n - 1
    embeds the same linear convolution, so the product is the same to
    rounding; e.g. the 102-voxel side of a 5 mm head model gets 210 instead of
    204 = 4 * 3 * 17."""
    m = 2 * n
    if _largest_prime_factor(m) <= 13:
        return m
    while _largest_prime_factor(m) > 7:
        m += 2
    return m


class Operator:
    """A device-resident JVIE operator: fN, material, DOF index and work buffers.

    Mirrors the state that MARIE keeps in the closures built by
    JVIE_Solver.m:271-278 (fN_gpu, Mr_gpu, Mc_gpu, idxS3).
    """

    def __init__(self, L, M, N, P=None, precisions=PREC_F64, device=0):
        lib = load()
        P = P or (circulant_side(L), circulant_side(M), circulant_side(N))
        self.L, self.M, self.N = L, M, N
        self.P = tuple(P)
        self.precisions = precisions
        h = ctypes.c_void_p()
        check(lib.marie_vie_create(ctypes.byref(h), L, M, N, P[0], P[1], P[2], precisions, device))
        self._h = h
        self.nidx = 0

    def close(self):
        if getattr(self, '_h', None):
            load().marie_vie_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        check(load().marie_vie_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    # -- operator ------------------------------------------------------------
    def set_fN(self, fN):
        Px, Py, Pz = self.P
        fN = np.asarray(fN)
        if fN.shape != (Px, Py, Pz, 6):
            raise ValueError('fN must be %s x 6' % (self.P,))
        buf = _cplx(fN)
        check(load().marie_vie_set_fN(self._h, _ptr(buf)))

    def set_fN_device(self, ptr):
        check(load().marie_vie_set_fN_device(self._h, ctypes.c_void_p(ptr)))

    def get_fN(self):
        Px, Py, Pz = self.P
        out = np.empty(Px * Py * Pz * 6, np.complex128)
        check(load().marie_vie_get_fN(self._h, _ptr(out)))
        return out.reshape(Px, Py, Pz, 6, order='F')

    def assemble_N(self, dx, ko, singular=None):
        if singular is not None:
            s = np.ascontiguousarray(np.asarray(singular, np.complex128))
            if s.shape != (5,):
                raise ValueError('singular must hold 5 complex values')
            check(load().marie_vie_assemble_N(self._h, float(dx), float(ko), _ptr(s)))
        else:
            check(load().marie_vie_assemble_N(self._h, float(dx), float(ko), None))

    def assemble_type(self, dx, ko, op=OP_N, type_='T', singular=None):
        """getOPERATORS' 'T' (Toeplitz generator, L x M x N x c) or 'C'
        (circulant, Px x Py x Pz x c) of the N (c = 6) or K (c = 3) operator,
        assembled on the device; the handle's operator is not changed."""
        nc = 6 if op == OP_N else 3
        shape = (self.L, self.M, self.N, nc) if type_ == 'T' else tuple(self.P) + (nc,)
        out = np.empty(int(np.prod(shape)), np.complex128)
        sv = None if singular is None else np.ascontiguousarray(np.asarray(singular, np.complex128))
        check(load().marie_vie_assemble_type(self._h, op, float(dx), float(ko), None if sv is None else _ptr(sv),
                                             0 if type_ == 'T' else 1, _ptr(out)))
        return out.reshape(shape, order='F')

    def set_fK(self, fK):
        Px, Py, Pz = self.P
        fK = np.asarray(fK)
        if fK.shape != (Px, Py, Pz, 3):
            raise ValueError('fK must be %s x 3' % (self.P,))
        buf = _cplx(fK)
        check(load().marie_vie_set_fK(self._h, _ptr(buf)))

    def get_fK(self):
        Px, Py, Pz = self.P
        out = np.empty(Px * Py * Pz * 3, np.complex128)
        check(load().marie_vie_get_fK(self._h, _ptr(out)))
        return out.reshape(Px, Py, Pz, 3, order='F')

    def assemble_K(self, dx, ko):
        check(load().marie_vie_assemble_K(self._h, float(dx), float(ko)))

    # -- material / index -----------------------------------------------------
    def set_material(self, Mr, Mc, Gram):
        shp = (self.L, self.M, self.N)
        Mr = np.broadcast_to(np.asarray(Mr, np.complex128), shp)
        Mc = np.broadcast_to(np.asarray(Mc, np.complex128), shp)
        a, b = _cplx(Mr), _cplx(Mc)
        check(load().marie_vie_set_material(self._h, _ptr(a), _ptr(b), float(Gram)))

    def set_index(self, idx, one_based=False):
        idx = np.ascontiguousarray(np.asarray(idx, np.int64).ravel())
        check(load().marie_vie_set_index(self._h, _ptr(idx), int(idx.size), 1 if one_based else 0))
        self.nidx = int(idx.size)

    # -- products ---------------------------------------------------------------
    def apply(self, x, op=OP_AN, transp=False):
        """Host-to-host product (complex128); x is nidx or nidx x nrhs (columns)."""
        x = np.asarray(x, np.complex128)
        if x.ndim not in (1, 2) or x.shape[0] != self.nidx:
            raise ValueError('x must be nidx = %d long (or nidx x nrhs); got shape %s' % (self.nidx, x.shape))
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xb = np.ascontiguousarray(x.T if x.ndim == 2 else x)
        y = np.empty_like(xb)
        check(load().marie_vie_apply(self._h, op, 1 if transp else 0, _ptr(xb), _ptr(y), nrhs))
        return y.T if x.ndim == 2 else y

    def apply_device(self, x_ptr, y_ptr, op=OP_AN, transp=False, prec=PREC_F64, nrhs=1):
        """Device product, enqueued on the handle's stream and not waited for:
        synchronise (or share the stream, set_stream) before reading y."""
        check(load().marie_vie_apply_device(self._h, op, 1 if transp else 0, prec, ctypes.c_void_p(x_ptr),
                                            ctypes.c_void_p(y_ptr), nrhs))

    def apply_device_noalloc(self, x_ptr, y_ptr, op=OP_AN, transp=False, prec=PREC_F64, nrhs=1):
        """apply_device without any allocation: the work buffers live in the
        library's per-device scratch array (noalloc_scratch_bytes()); calls
        that overlap on one GPU race."""
        check(load().marie_vie_apply_device_noalloc(self._h, op, 1 if transp else 0, prec, ctypes.c_void_p(x_ptr),
                                                    ctypes.c_void_p(y_ptr), nrhs))

    def gmres_device(self, b_ptr, x_ptr, restart=50, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4,
                     resvec_cap=20000):
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_device(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr), restart, tol,
                                            maxit, mode, inner_tol, ctypes.byref(flag), ctypes.byref(relres), it,
                                            resvec.ctypes.data_as(_DP), resvec_cap, ctypes.byref(nres)))
        return flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def gmres_adjoint_device(self, g_ptr, y_ptr, restart=50, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4,
                             resvec_cap=20000):
        """GMRES on A^H y = g (device buffers); returns (flag, relres, [inner, outer], resvec)."""
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_adjoint_device(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(y_ptr), restart,
                                                    tol, maxit, mode, inner_tol, ctypes.byref(flag),
                                                    ctypes.byref(relres), it, resvec.ctypes.data_as(_DP), resvec_cap,
                                                    ctypes.byref(nres)))
        return flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def material_gradient_device(self, lam_ptr, x_ptr, einc_ptr, omega_eo, grad_eps_ptr, grad_sigma_ptr):
        """Per-voxel dJ/deps_r, dJ/dsigma into device buffers (L*M*N doubles each)."""
        check(load().marie_vie_material_gradient_device(self._h, ctypes.c_void_p(lam_ptr), ctypes.c_void_p(x_ptr),
                                                        ctypes.c_void_p(einc_ptr), omega_eo,
                                                        ctypes.c_void_p(grad_eps_ptr),
                                                        ctypes.c_void_p(grad_sigma_ptr)))

    def block_gmres_device(self, B_ptr, X_ptr, nrhs, restart=20, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4,
                           use_graph=False, cgs_passes=1, resvec_cap=20000, columns=False, device_steps=False):
        """Block GMRES on device buffers (columns=True: the columns one at a
        time with GMRES in `mode` 0-3; device_steps=True: every block step on
        the device, a HIP graph per step with use_graph); returns (flags,
        relres, [steps, cycles], resvec)."""
        flag = (ctypes.c_int * nrhs)()
        relres = (ctypes.c_double * nrhs)()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_block_gmres_device(self._h, ctypes.c_void_p(B_ptr), ctypes.c_void_p(X_ptr), nrhs,
                                                  restart, tol, maxit, mode, inner_tol,
                                                  _block_options(use_graph, cgs_passes, columns, device_steps),
                                                  flag, relres, it, resvec.ctypes.data_as(_DP), resvec_cap,
                                                  ctypes.byref(nres)))
        return (np.array(flag[:]), np.array(relres[:]), [it[0], it[1]],
                resvec[:min(nres.value, resvec_cap)].copy())

    def block_gmres(self, B, X0=None, restart=20, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4, use_graph=False,
                    cgs_passes=1, resvec_cap=20000, device_steps=False):
        """Host-buffer block GMRES; B is nidx x nrhs. Returns (X, flags, relres, [steps, cycles], resvec)."""
        B = np.asarray(B, np.complex128)
        if B.ndim == 1:
            B = B[:, None]
        if B.ndim != 2 or B.shape[0] != self.nidx:
            raise ValueError('B must be nidx = %d x nrhs; got shape %s' % (self.nidx, B.shape))
        nrhs = B.shape[1]
        Bc = np.ascontiguousarray(B.T)
        if X0 is not None and np.asarray(X0).size != B.size:
            raise ValueError('X0 must hold nidx x nrhs = %d entries' % B.size)
        Xc = (np.zeros_like(Bc) if X0 is None else
              np.ascontiguousarray(np.asarray(X0, np.complex128).reshape(B.shape).T).copy())
        flag = (ctypes.c_int * nrhs)()
        relres = (ctypes.c_double * nrhs)()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_block_gmres(self._h, _ptr(Bc), _ptr(Xc), nrhs, restart, tol, maxit, mode, inner_tol,
                                           _block_options(use_graph, cgs_passes, False, device_steps), flag,
                                           relres, it,
                                           resvec.ctypes.data_as(_DP),
                                           resvec_cap, ctypes.byref(nres)))
        return (Xc.T.copy(), np.array(flag[:]), np.array(relres[:]), [it[0], it[1]],
                resvec[:min(nres.value, resvec_cap)].copy())

    def gmres(self, b, x0=None, restart=50, tol=1e-6, maxit=200, mode=0, inner_tol=1e-4, resvec_cap=20000):
        """Host-buffer GMRES; returns (x, flag, relres, [inner, outer], resvec)."""
        b, x = _vec_arg(b, x0, self.nidx)
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres(self._h, _ptr(b), _ptr(x), restart, tol, maxit, mode, inner_tol,
                                     ctypes.byref(flag), ctypes.byref(relres), it, resvec.ctypes.data_as(_DP),
                                     resvec_cap, ctypes.byref(nres)))
        return x, flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def gmres_precond(self, b, x0=None, restart=50, tol=1e-6, maxit=200, side=PRECOND_RIGHT, resvec_cap=20000):
        """Host-buffer preconditioned GMRES (pgmres with R1 or L1 = the JVIE
        preconditioner); returns (x, flag, relres, [inner, outer], resvec)."""
        b, x = _vec_arg(b, x0, self.nidx)
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_precond(self._h, _ptr(b), _ptr(x), restart, tol, maxit, side,
                                             ctypes.byref(flag), ctypes.byref(relres), it,
                                             resvec.ctypes.data_as(_DP), resvec_cap, ctypes.byref(nres)))
        return x, flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def gmres_precond_device(self, b_ptr, x_ptr, restart=50, tol=1e-6, maxit=200, side=PRECOND_RIGHT,
                             resvec_cap=20000):
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_precond_device(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr),
                                                    restart, tol, maxit, side, ctypes.byref(flag),
                                                    ctypes.byref(relres), it, resvec.ctypes.data_as(_DP),
                                                    resvec_cap, ctypes.byref(nres)))
        return flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def gmres_dr(self, b, x0=None, restart=50, tol=1e-6, maxit=200, ritz=5, side=PRECOND_NONE,
                 resvec_cap=20000):
        """Host-buffer GMRES with deflated restarts (pgmresDR); returns
        (x, flag, relres, [inner, outer], resvec)."""
        b, x = _vec_arg(b, x0, self.nidx)
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_dr(self._h, _ptr(b), _ptr(x), restart, tol, maxit, ritz, side,
                                        ctypes.byref(flag), ctypes.byref(relres), it, resvec.ctypes.data_as(_DP),
                                        resvec_cap, ctypes.byref(nres)))
        return x, flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def gmres_dr_device(self, b_ptr, x_ptr, restart=50, tol=1e-6, maxit=200, ritz=5, side=PRECOND_NONE,
                        resvec_cap=20000):
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = (ctypes.c_int * 2)()
        resvec = np.zeros(resvec_cap)
        nres = ctypes.c_int()
        check(load().marie_vie_gmres_dr_device(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr), restart, tol,
                                               maxit, ritz, side, ctypes.byref(flag), ctypes.byref(relres), it,
                                               resvec.ctypes.data_as(_DP), resvec_cap, ctypes.byref(nres)))
        return flag.value, relres.value, [it[0], it[1]], resvec[:min(nres.value, resvec_cap)].copy()

    def bicgstab(self, b, x0=None, tol=1e-6, maxit=1000, precond=False):
        """Host-buffer BiCGStab (MATLAB bicgstab semantics); returns
        (x, flag, relres, iter, resvec) with iter in half steps / 2 and resvec
        the absolute residual norms of every half step."""
        b, x = _vec_arg(b, x0, self.nidx)
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = ctypes.c_double()
        cap = 2 * maxit + 1
        resvec = np.zeros(cap)
        nres = ctypes.c_int()
        check(load().marie_vie_bicgstab(self._h, _ptr(b), _ptr(x), tol, maxit, 1 if precond else 0,
                                        ctypes.byref(flag), ctypes.byref(relres), ctypes.byref(it),
                                        resvec.ctypes.data_as(_DP), cap, ctypes.byref(nres)))
        return x, flag.value, relres.value, it.value, resvec[:min(nres.value, cap)].copy()

    def bicgstab_device(self, b_ptr, x_ptr, tol=1e-6, maxit=1000, precond=False):
        flag = ctypes.c_int()
        relres = ctypes.c_double()
        it = ctypes.c_double()
        cap = 2 * maxit + 1
        resvec = np.zeros(cap)
        nres = ctypes.c_int()
        check(load().marie_vie_bicgstab_device(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr), tol, maxit,
                                               1 if precond else 0, ctypes.byref(flag), ctypes.byref(relres),
                                               ctypes.byref(it), resvec.ctypes.data_as(_DP), cap,
                                               ctypes.byref(nres)))
        return flag.value, relres.value, it.value, resvec[:min(nres.value, cap)].copy()

    # -- tiles ------------------------------------------------------------------
    def autotune(self, prec=PREC_F64, iters=5):
        """Pick the fastest variant per kernel; returns (cfg6, best_ms5)."""
        em = (ctypes.c_int * 6)()
        ms = (ctypes.c_double * 5)()
        check(load().marie_vie_autotune(self._h, prec, iters, em, ms))
        return list(em), list(ms)

    def set_tiles(self, cfg, prec=PREC_F64):
        """cfg: five per-kernel variants, optionally a sixth z-grouping flag."""
        cfg = list(cfg) + [0] * (6 - len(cfg))
        em = (ctypes.c_int * 6)(*cfg)
        check(load().marie_vie_set_tiles(self._h, prec, em))

    # -- timing -------------------------------------------------------------------
    def enable_timing(self, on=True):
        check(load().marie_vie_enable_timing(self._h, 1 if on else 0))

    def timing(self):
        ms = (ctypes.c_double * 5)()
        calls = ctypes.c_int64()
        check(load().marie_vie_get_timing(self._h, ms, ctypes.byref(calls)))
        return list(ms), calls.value


def noalloc_scratch_bytes():
    """Size of the per-device scratch array of apply_device_noalloc."""
    return int(load().marie_vie_noalloc_scratch_bytes())


def singular_terms(dx, ko):
    """The library's five self-term integrals (I_ST, I_EAc, I_EAo, I_VAc, I_VAo)."""
    out = np.zeros(5, np.complex128)
    check(load().marie_vie_singular_terms(float(dx), float(ko), _ptr(out)))
    return out


def singular_terms_directfn(dx, ko, Np=20):
    """The five self-terms by DIRECTFN with Np Gauss points per direction, as
    the reference's default method computes them (singular_st/ea/va.m)."""
    out = np.zeros(5, np.complex128)
    check(load().marie_vie_singular_terms_directfn(float(dx), float(ko), int(Np), _ptr(out)))
    return out
