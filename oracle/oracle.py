"""ctypes binding of the CPU oracle (liboracle_evr.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def lib():
    global _LIB
    if _LIB is None:
        # ORACLE_LIB: the same source built for one host (the CPU baseline's
        # -march=native copy)
        path = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle_evr.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_sg4_create.restype = ctypes.c_void_p
        L.orc_sg4_create.argtypes = [ctypes.c_int] * 7
        L.orc_sg4_create_modes.restype = ctypes.c_void_p
        L.orc_sg4_create_modes.argtypes = [ctypes.c_int] * 3 + [_ip] * 4
        L.orc_sg4_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sg4_sizes.argtypes = [ctypes.c_void_p, _llp]
        L.orc_sg4_terms.argtypes = [ctypes.c_void_p, _ip, _dp, _llp, _llp]
        L.orc_sg4_packed.argtypes = [ctypes.c_void_p, _ip]
        L.orc_sg4_map.argtypes = [ctypes.c_void_p, _ip]
        L.orc_sg4_grid.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sg4_1d.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                 _dp, _dp, _dp, _dp, _dp]
        L.orc_sg4_oppsi.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, _dp,
                                    ctypes.c_int]
        L.orc_sg4_oppsi_range.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp,
                                          _dp, ctypes.c_int, ctypes.c_longlong,
                                          ctypes.c_longlong]
        L.orc_sg4_oppsi_terms.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp,
                                          _dp, ctypes.c_int, _llp, ctypes.c_longlong]
        L.orc_sg4_b2g.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_sg4_g2b.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_sg4_deriv.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_dp_deriv.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_herm_quadrature.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_herm_basis.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_herm_dngg.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_dp_create.restype = ctypes.c_void_p
        L.orc_dp_create.argtypes = [ctypes.c_int, _ip, _ip]
        L.orc_dp_destroy.argtypes = [ctypes.c_void_p]
        L.orc_dp_sizes.argtypes = [ctypes.c_void_p, _llp]
        L.orc_dp_grid.argtypes = [ctypes.c_void_p, _dp]
        L.orc_dp_oppsi.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_dp_b2g.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_dp_g2b.argtypes = [ctypes.c_void_p, _dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip)


class _Op(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("V", _dp), ("G0", _dp), ("Ga", _dp), ("F2", _dp),
                ("F1", _dp), ("Jac", _dp), ("sqRhoOVERJac", _dp), ("Gq", _dp), ("Vpoly", _dp),
                ("F2q", _dp), ("F1q", _dp)]


def poly_coefs(v2, c3, d22, e4):
    """Coefficient array of the oracle's per-point polynomial potential:
    v2[D], c3 as (i*D + j)*D + k, d22 as i*D + j, e4[D]."""
    D = len(v2)
    return np.concatenate([np.asarray(v2, float), np.asarray(c3, float).reshape(D ** 3),
                           np.asarray(d22, float).reshape(D * D), np.asarray(e4, float)])


class Op:
    """Operator description (type 0 / 1 / 10), arrays kept alive here."""

    def __init__(self, type, V=None, G0=None, Ga=None, F2=None, F1=None, Jac=None, s=None, Gq=None,
                 Vpoly=None, F2q=None, F1q=None):
        """Gq: SG4 type 10 G per grid point, (sum_nq, D, D) with Gq[g, j, i] = GGiq(g, j, i).
        Vpoly: poly_coefs(...) of a polynomial potential evaluated per grid
        point (SG4, in place of V).  F2q (sum_nq, D, D), F1q (sum_nq, D): SG4
        type 1 operator grids per point, F2q[g, i, j] multiplying d_i d_j."""
        def f(a):
            return None if a is None else np.asarray(a, dtype=np.float64).ravel(order="F").copy()
        self.arrays = dict(V=f(V), G0=f(G0), Ga=f(Ga), F2=f(F2), F1=f(F1), Jac=f(Jac), s=f(s), Gq=f(Gq),
                           Vpoly=f(Vpoly), F2q=f(F2q), F1q=f(F1q))
        a = self.arrays
        self.st = _Op(type, _d(a["V"]), _d(a["G0"]), _d(a["Ga"]), _d(a["F2"]), _d(a["F1"]),
                      _d(a["Jac"]), _d(a["s"]), _d(a["Gq"]), _d(a["Vpoly"]), _d(a["F2q"]), _d(a["F1q"]))


def herm_quadrature(nq):
    x = np.zeros(nq)
    w = np.zeros(nq)
    lib().orc_herm_quadrature(nq, _d(x), _d(w))
    return x, w


def herm_basis(nb, x):
    nq = len(x)
    x = np.ascontiguousarray(x, dtype=np.float64)
    B = np.zeros((nq, nb), order="F")
    B1 = np.zeros((nq, nb), order="F")
    B2 = np.zeros((nq, nb), order="F")
    lib().orc_herm_basis(nb, nq, _d(x), _d(B), _d(B1), _d(B2))
    return B, B1, B2


def herm_dngg(B, B1, B2):
    nq, nb = B.shape
    B = np.asfortranarray(B); B1 = np.asfortranarray(B1); B2 = np.asfortranarray(B2)
    D0 = np.zeros((nq, nq), order="F")
    D1 = np.zeros((nq, nq), order="F")
    D2 = np.zeros((nq, nq), order="F")
    lib().orc_herm_dngg(nb, nq, _d(B), _d(B1), _d(B2), _d(D0), _d(D1), _d(D2))
    return D0, D1, D2


class SG4:
    """Smolyak SG4 basis of the oracle: nb(l)=nbA+nbB*l, nq(l)=nqA+nqB*l."""

    def __init__(self, D, LG, LB=None, nbA=1, nbB=2, nqA=2, nqB=2):
        """nbA..nqB: integers, or sequences of D integers (a law per mode)."""
        LB = LG if LB is None else LB
        if all(np.ndim(v) == 0 for v in (nbA, nbB, nqA, nqB)):
            self.h = lib().orc_sg4_create(D, LG, LB, nbA, nbB, nqA, nqB)
        else:
            arr = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.int32), (D,)))
                   for v in (nbA, nbB, nqA, nqB)]
            self.h = lib().orc_sg4_create_modes(D, LG, LB, *[_i(a) for a in arr])
        if not self.h:
            raise ValueError("orc_sg4_create failed")
        z = np.zeros(8, dtype=np.int64)
        lib().orc_sg4_sizes(self.h, z.ctypes.data_as(_llp))
        (self.nterms, self.nb_packed, self.sum_nb, self.sum_nq, self.max_nb, self.max_nq,
         self.D, self.LG) = [int(v) for v in z]
        self.LB = LB

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sg4_destroy(self.h)
            self.h = None

    def terms(self):
        lv = np.zeros((self.nterms, self.D), dtype=np.int32)
        w = np.zeros(self.nterms)
        onb = np.zeros(self.nterms + 1, dtype=np.int64)
        onq = np.zeros(self.nterms + 1, dtype=np.int64)
        lib().orc_sg4_terms(self.h, _i(lv), _d(w), onb.ctypes.data_as(_llp),
                            onq.ctypes.data_as(_llp))
        return lv, w, onb, onq

    def packed(self):
        t = np.zeros((self.nb_packed, self.D), dtype=np.int32)
        lib().orc_sg4_packed(self.h, _i(t))
        return t

    def map(self):
        m = np.zeros(self.sum_nb, dtype=np.int32)
        lib().orc_sg4_map(self.h, _i(m))
        return m

    def grid(self):
        q = np.zeros((self.sum_nq, self.D))
        lib().orc_sg4_grid(self.h, _d(q))
        return q

    def one_d(self, d, l):
        nb = ctypes.c_int(); nq = ctypes.c_int()
        lib().orc_sg4_1d(self.h, d, l, ctypes.byref(nb), ctypes.byref(nq), None, None, None,
                         None, None)
        nb, nq = nb.value, nq.value
        x = np.zeros(nq)
        B = np.zeros((nq, nb), order="F"); W = np.zeros((nb, nq), order="F")
        D1 = np.zeros((nq, nq), order="F"); D2 = np.zeros((nq, nq), order="F")
        lib().orc_sg4_1d(self.h, d, l, ctypes.byref(ctypes.c_int()), ctypes.byref(ctypes.c_int()),
                         _d(x), _d(B), _d(W), _d(D1), _d(D2))
        return dict(nb=nb, nq=nq, x=x, B=B, W=W, D1=D1, D2=D2)

    def oppsi(self, op, psi, nthreads=0, t0=None, t1=None):
        psi = np.asfortranarray(psi, dtype=np.float64)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1, order="F")
        assert psi.shape[0] == self.nb_packed
        out = np.zeros_like(psi, order="F")
        if t0 is None and t1 is None:
            rc = lib().orc_sg4_oppsi(self.h, ctypes.addressof(op.st), psi.shape[1], _d(psi),
                                     _d(out), nthreads)
        else:
            rc = lib().orc_sg4_oppsi_range(self.h, ctypes.addressof(op.st), psi.shape[1],
                                           _d(psi), _d(out), nthreads, t0 or 0,
                                           self.nterms if t1 is None else t1)
        if rc != 0:
            raise RuntimeError("orc_sg4_oppsi failed: %d" % rc)
        return out

    def oppsi_terms(self, op, psi, terms, nthreads=0):
        psi = np.asfortranarray(psi, dtype=np.float64)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        terms = np.ascontiguousarray(terms, dtype=np.int64)
        rc = lib().orc_sg4_oppsi_terms(self.h, ctypes.addressof(op.st), psi.shape[1], _d(psi),
                                       _d(out), nthreads, terms.ctypes.data_as(_llp), len(terms))
        if rc != 0:
            raise RuntimeError("orc_sg4_oppsi_terms failed: %d" % rc)
        return out

    def b2g(self, psiB):
        psiB = np.ascontiguousarray(psiB, dtype=np.float64)
        g = np.zeros(self.sum_nq)
        lib().orc_sg4_b2g(self.h, _d(psiB), _d(g))
        return g

    def g2b(self, psiG):
        psiG = np.ascontiguousarray(psiG, dtype=np.float64)
        b = np.zeros(self.nb_packed)
        lib().orc_sg4_g2b(self.h, _d(psiG), _d(b))
        return b

    def deriv(self, psiG, i, j):
        """Grid derivative on every term grid; i, j 0-based modes, -1 for none."""
        psiG = np.ascontiguousarray(psiG, dtype=np.float64)
        out = np.zeros(self.sum_nq)
        if lib().orc_sg4_deriv(self.h, i, j, _d(psiG), _d(out)) != 0:
            raise ValueError("bad mode")
        return out


class DP:
    """Direct-product Hermite basis of the oracle (last dimension fastest)."""

    def __init__(self, nb, nq):
        self.D = len(nb)
        nb = np.asarray(nb, dtype=np.int32); nq = np.asarray(nq, dtype=np.int32)
        self.h = lib().orc_dp_create(self.D, _i(nb), _i(nq))
        z = np.zeros(2, dtype=np.int64)
        lib().orc_dp_sizes(self.h, z.ctypes.data_as(_llp))
        self.nb_tot, self.nq_tot = int(z[0]), int(z[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dp_destroy(self.h)
            self.h = None

    def grid(self):
        q = np.zeros((self.nq_tot, self.D))
        lib().orc_dp_grid(self.h, _d(q))
        return q

    def oppsi(self, op, psi):
        psi = np.asfortranarray(psi, dtype=np.float64)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        lib().orc_dp_oppsi(self.h, ctypes.addressof(op.st), psi.shape[1], _d(psi), _d(out))
        return out

    def b2g(self, psiB):
        psiB = np.ascontiguousarray(psiB, dtype=np.float64)
        g = np.zeros(self.nq_tot)
        lib().orc_dp_b2g(self.h, _d(psiB), _d(g))
        return g

    def g2b(self, psiG):
        psiG = np.ascontiguousarray(psiG, dtype=np.float64)
        b = np.zeros(self.nb_tot)
        lib().orc_dp_g2b(self.h, _d(psiG), _d(b))
        return b

    def deriv(self, psiG, i, j):
        psiG = np.ascontiguousarray(psiG, dtype=np.float64)
        out = np.zeros(self.nq_tot)
        if lib().orc_dp_deriv(self.h, i, j, _d(psiG), _d(out)) != 0:
            raise ValueError("bad mode")
        return out
