"""Host-side mirror of ElVibRot's operator interface over the HIP engine.

Names follow the reference: a Plan's `TabOpPsi` is sub_TabOpPsi
(Source_ElVibRot/sub_Operator/sub_OpPsi.f90:701), `BasisRep_TO_GridRep` /
`GridRep_TO_BasisRep` are sub_PsiBasisRep_TO_GridRep /
sub_PsiGridRep_TO_BasisRep (Source_ElVibRot/sub_WP/sub_module_psi_B_TO_G.f90:60,
178).  Everything runs in libevrhip.so; there is no CPU fallback: when the
library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

from . import hamiltonian  # noqa: F401
from ._lib import (BasisDesc, ControlInfo, ControlOpts, DavidsonInfo, DavidsonOpts, OpDesc, PropaInfo,
                   PropaOpts, check, load)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

FP64 = 0
FP32 = 1
DEVICE_PTRS = 2
MATRIX_FREE = 4
AUTO_PATH = 8
NOALLOC_SCRATCH_BYTES = 64 << 20   # EVR_NOALLOC_SCRATCH_BYTES of evr_hip.h
ASM_FP64 = 1
ASM_FP32 = 2
CHEBY = 1
SIL = 2
TAYLOR = 3
RK4 = 4


def _to_pairs(psi):
    """Complex (nb, nwp) -> real nb x 2nwp column-major, columns (Re, Im) per wavepacket."""
    psi = np.asarray(psi, dtype=np.complex128)
    if psi.ndim == 1:
        psi = psi[:, None]
    r = np.empty((psi.shape[0], 2 * psi.shape[1]), order="F")
    r[:, 0::2] = psi.real
    r[:, 1::2] = psi.imag
    return r


def _from_pairs(r):
    return r[:, 0::2] + 1j * r[:, 1::2]


def control_field(amp, T, nsteps):
    """Mid-step field values (nsteps, nf) of E(t) = sin^2(pi t/T) sum_k a_k sin(k pi t/T)."""
    amp = np.asfortranarray(np.asarray(amp, dtype=np.float64).reshape(np.shape(amp)[0], -1))
    nk, nf = amp.shape
    o = ControlOpts(T, nsteps, nk, 0, 3, 0.0, 0.0, 0.0)
    out = np.zeros((nsteps, nf))
    check(load().evr_control_field(ctypes.byref(o), nf, _d(amp), _d(out)))
    return out


def control_build_field(grad_E, T, alpha, Tenv=0.0, field_old=None):
    """build_field: the Zhu-Rabitz field (or, with field_old, the Krotov update)
    at the mid-steps from dJ/dE_n (nsteps, nf)."""
    g = np.ascontiguousarray(grad_E, dtype=np.float64)
    nsteps, nf = g.shape
    o = ControlOpts(T, nsteps, 1, 0, 3, 1e-12, 0.0, 0.0)
    old = None if field_old is None else np.ascontiguousarray(field_old, dtype=np.float64)
    out = np.zeros((nsteps, nf))
    check(load().evr_control_build_field(ctypes.byref(o), nf, _d(g), alpha, Tenv, 0 if old is None else 1,
                                         _d(old), _d(out)))
    return out


def control_field_sub(amp, T, nsteps, method, nsub):
    """The control fields sampled for the TAYLOR (derivatives at t_n) or RK4
    (t_n, t_n + dt/2, t_n + dt) integrators: (nsteps, nsub, nf)."""
    amp = np.asfortranarray(amp, dtype=np.float64)
    if amp.ndim == 1:
        amp = amp.reshape(-1, 1, order="F")
    nk, nf = amp.shape
    o = ControlOpts(T, nsteps, nk, 0, 3, 1e-12, 0.0, 0.0)
    out = np.zeros((nsteps, nsub, nf))
    check(load().evr_control_field_sub(ctypes.byref(o), nf, _d(amp), method, nsub, _d(out)))
    return out


def save_restart(path, T, psi):
    """SaveWP_restart: time T and complex wavepackets psi (size_WP,) or
    (size_WP, nwp)."""
    psi = np.asarray(psi)
    if psi.ndim == 1:
        psi = psi[:, None]
    r = _to_pairs(psi)
    check(load().evr_propa_save_restart(T, psi.shape[1], psi.shape[0], _d(r), os.fsencode(path)))


def read_restart(path, nwp, size_wp):
    """ReadWP_restart: (T, psi (size_WP, nwp) complex, found)."""
    r = np.zeros((size_wp, 2 * nwp), order="F")
    T = ctypes.c_double()
    found = ctypes.c_int()
    check(load().evr_propa_read_restart(os.fsencode(path), nwp, size_wp, _d(r), ctypes.byref(T),
                                        ctypes.byref(found)))
    return T.value, _from_pairs(r), bool(found.value)


def bessel_j(n, r):
    """J_0(r) .. J_n(r) as computed for the Chebyshev coefficients."""
    out = np.zeros(n + 1)
    check(load().evr_bessel_j(n, float(r), _d(out)))
    return out


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Basis:
    """An SG4 or direct-product Hermite basis description (owned by the library)."""

    def __init__(self, ptr):
        self.ptr = ptr
        d = ptr.contents
        self.D, self.nlev, self.nterms, self.nb_packed = d.D, d.nlev, d.nterms, d.nb_packed
        z = (ctypes.c_int64 * 4)()
        check(load().evr_basis_desc_sizes(ptr, z))
        self.sum_nb, self.sum_nq, self.max_nb, self.max_nq = list(z)

    @classmethod
    def sg4(cls, D, LG, LB=None, nbA=1, nbB=2, nqA=2, nqB=2):
        LB = LG if LB is None else LB
        p = ctypes.POINTER(BasisDesc)()
        check(load().evr_sg4_hermite_desc(D, LG, LB, nbA, nbB, nqA, nqB, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def sg4_modes(cls, D, LG, nbA, nbB, nqA, nqB, LB=None):
        """SG4 Hermite basis with a law per mode: nb_d(l) = nbA[d] + nbB[d] l,
        nq_d(l) = nqA[d] + nqB[d] l (sequences of length D)."""
        LB = LG if LB is None else LB
        arr = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.int32), (D,))) for v in
               (nbA, nbB, nqA, nqB)]
        p = ctypes.POINTER(BasisDesc)()
        check(load().evr_sg4_hermite_desc_modes(D, LG, LB, *[a.ctypes.data_as(_ip) for a in arr],
                                                ctypes.byref(p)))
        return cls(p)

    @classmethod
    def dp(cls, nb, nq):
        nb = np.asarray(nb, dtype=np.int32)
        nq = np.asarray(nq, dtype=np.int32)
        p = ctypes.POINTER(BasisDesc)()
        check(load().evr_dp_hermite_desc(len(nb), nb.ctypes.data_as(_ip), nq.ctypes.data_as(_ip),
                                         ctypes.byref(p)))
        return cls(p)

    def grid(self):
        q = np.zeros((self.sum_nq, self.D))
        check(load().evr_basis_desc_grid(self.ptr, _d(q)))
        return q

    def term_levels(self):
        d = self.ptr.contents
        return np.ctypeslib.as_array(d.term_levels, shape=(self.nterms, self.D)).copy()

    def weights(self):
        return np.ctypeslib.as_array(self.ptr.contents.weights, shape=(self.nterms,)).copy()

    def map(self):
        return np.ctypeslib.as_array(self.ptr.contents.map, shape=(self.sum_nb,)).copy()

    def save_psi(self, path, psi, nb0=1, version=0, formatted=True):
        """Write wavepackets (nb0*nb_packed,) or (nb0*nb_packed, nwp), real or
        complex, in the reference's save-file format (sub_save_psi)."""
        psi = np.asarray(psi)
        cplx = np.iscomplexobj(psi)
        if psi.ndim == 1:
            psi = psi[:, None]
        arr = _to_pairs(psi) if cplx else np.asfortranarray(psi, dtype=np.float64)
        check(load().evr_psi_save(self.ptr, nb0, psi.shape[1], 1 if cplx else 0, _d(arr), version,
                                  1 if formatted else 0, os.fsencode(path)))

    def read_psi(self, path, nb0=1, cplx=True, max_wp=1):
        """Wavepackets of a save file projected onto this basis:
        (nb0*nb_packed, nwp) complex (or real) array."""
        n = self.nb_packed * nb0
        arr = np.zeros((n, (2 if cplx else 1) * max_wp), order="F")
        got = ctypes.c_int()
        check(load().evr_psi_read(self.ptr, nb0, 1 if cplx else 0, max_wp, _d(arr), ctypes.byref(got),
                                  os.fsencode(path)))
        out = _from_pairs(arr) if cplx else arr
        return out[:, :got.value]

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                load().evr_basis_desc_free(self.ptr)
        except Exception:
            pass
        self.ptr = None


class Plan:
    """Device plan of one operator on one basis (para_Op + basis on the GPU)."""

    def __init__(self, basis, type=10, V=None, G0=None, Ga=None, F2=None, F1=None, max_nvec=1,
                 hermitian=False):
        self.basis = basis
        f = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, float).ravel(order="F"))
        self._keep = [f(V), f(G0), f(Ga), f(F2), f(F1)]
        op = OpDesc(type, *[_d(a) for a in self._keep], 1 if hermitian else 0)
        h = ctypes.c_void_p()
        check(load().evr_plan_create(basis.ptr, ctypes.byref(op), max_nvec, ctypes.byref(h)))
        self.h = h
        self.max_nvec = max_nvec
        z = (ctypes.c_int64 * 7)()
        check(load().evr_plan_sizes(h, z))
        self.nb_packed = z[0]

    def set_potential(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        check(load().evr_plan_set_potential(self.h, _d(V)))

    def set_keo_grid(self, Gq, Jac=None, s=None):
        """Type 10 with G, J and s = sqrt(rho/J) per grid point: Gq is
        (sum_nq, D, D) with Gq[g, j, i] = GGiq(g, j, i) of the reference."""
        G = np.asarray(Gq, dtype=np.float64).ravel(order="F").copy()
        J = None if Jac is None else np.ascontiguousarray(Jac, dtype=np.float64)
        S = None if s is None else np.ascontiguousarray(s, dtype=np.float64)
        check(load().evr_plan_set_keo_grid(self.h, _d(G), _d(J), _d(S)))

    def set_type1_grid(self, F2q, F1q=None):
        """Type 1 with the operator grids per point: F2q (sum_nq, D, D) with
        F2q[g, i, j] multiplying d_i d_j, F1q (sum_nq, D) multiplying d_i."""
        F2 = np.asarray(F2q, dtype=np.float64).ravel(order="F").copy()
        F1 = None if F1q is None else np.asarray(F1q, dtype=np.float64).ravel(order="F").copy()
        check(load().evr_plan_set_type1_grid(self.h, _d(F2), _d(F1)))

    def gen_potential_poly(self, ham):
        D = ham.D
        c3 = np.ascontiguousarray(ham.c3.reshape(D * D * D))   # (i*D + j)*D + k
        d22 = np.ascontiguousarray(ham.d22.reshape(D * D))     # i*D + j
        check(load().evr_plan_gen_potential_poly(self.h, _d(np.ascontiguousarray(ham.v2)), _d(c3),
                                                 _d(d22), _d(np.ascontiguousarray(ham.e4))))

    def TabOpPsi(self, psi, flags=FP64):
        """H|psi> for a block of vectors (nb_packed x nvec), host arrays."""
        psi = np.asfortranarray(psi, dtype=np.float64)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        check(load().evr_oppsi(self.h, psi.shape[1], psi.ctypes.data, out.ctypes.data, flags))
        return out

    def TabOpPsi_device(self, psi_ptr, hpsi_ptr, nvec, flags=FP64):
        """H|psi> on device buffers (column-major), queued on the plan stream."""
        check(load().evr_oppsi(self.h, nvec, ctypes.c_void_p(psi_ptr), ctypes.c_void_p(hpsi_ptr),
                               flags | DEVICE_PTRS))

    def TabOpPsi_noalloc_device(self, psi_ptr, hpsi_ptr, nvec, flags=FP64):
        """evr_oppsi_noalloc: the same H|psi> with the work memory in the
        library's static per-GPU scratch (no device allocation)."""
        check(load().evr_oppsi_noalloc(self.h, nvec, ctypes.c_void_p(psi_ptr), ctypes.c_void_p(hpsi_ptr),
                                       flags | DEVICE_PTRS))

    def stream(self):
        return load().evr_plan_stream(self.h)

    def autotune(self, nvec, flags=FP64, reps=3):
        """Per-term tile autotuner: (chosen class per bucket, ms static, ms tuned)."""
        ch = (ctypes.c_int * 8)()
        a, b = ctypes.c_double(), ctypes.c_double()
        check(load().evr_plan_autotune(self.h, nvec, flags, reps, ch, ctypes.byref(a), ctypes.byref(b)))
        return list(ch), a.value, b.value

    def set_profiling(self, on=True):
        check(load().evr_plan_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self):
        """[(name, ms, flops, launches)] accumulated while profiling."""
        mx = 16
        names = ctypes.create_string_buffer(64 * mx)
        ms = np.zeros(mx); fl = np.zeros(mx); la = np.zeros(mx, dtype=np.int64)
        n = ctypes.c_int()
        check(load().evr_plan_kernel_stats(self.h, mx, names, _d(ms), _d(fl),
                                           la.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           ctypes.byref(n)))
        out = []
        for k in range(n.value):
            nm = names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode()
            out.append((nm, ms[k], fl[k], int(la[k])))
        return out

    def assemble(self, which=ASM_FP64):
        """Build the per-term dense operators (ASM_FP64 | ASM_FP32) used by TabOpPsi."""
        check(load().evr_plan_assemble(self.h, which))

    def assembled(self):
        """(mask of assembled precisions, bytes held)."""
        b = ctypes.c_double()
        m = load().evr_plan_assembled(self.h, ctypes.byref(b))
        if m < 0:
            check(m)
        return m, b.value

    def release_assembled(self):
        check(load().evr_plan_release_assembled(self.h))

    def diagonal(self):
        """Diagonal of H in the packed basis."""
        d = np.zeros(self.nb_packed)
        check(load().evr_plan_diagonal(self.h, _d(d), 0))
        return d

    def davidson(self, nev, block=16, max_sub=0, max_it=200, conv_resi=1e-8, conv_ene=1e-9,
                 precond=True, diag=None, guess=None, precision=FP64, refine_resi=1e-3,
                 want_vectors=False, verbose=False, guess_file=None, save_file=None, graph=False):
        """Lowest nev levels by block Davidson (sub_propagation_Davidson).
        guess_file / save_file: wavepacket files (sub_save_psi formats) read as
        start vectors / written with the nev Ritz vectors."""
        keep = []
        dg = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        gs = None if guess is None else np.asfortranarray(guess, dtype=np.float64)
        keep += [dg, gs]
        o = DavidsonOpts(nev, block, max_sub, max_it, conv_resi, conv_ene, 1 if precond else 0,
                         _d(dg), _d(gs), 0 if gs is None else gs.shape[1], precision, refine_resi,
                         min(block, self.max_nvec), int(verbose),
                         ctypes.cast(self.basis.ptr, ctypes.c_void_p) if (guess_file or save_file) else None,
                         None if guess_file is None else os.fsencode(guess_file),
                         None if save_file is None else os.fsencode(save_file), 1 if graph else 0)
        ev = np.zeros(nev)
        vec = np.zeros((self.nb_packed, nev), order="F") if want_vectors else None
        info = DavidsonInfo()
        check(load().evr_davidson(self.h, ctypes.byref(o), _d(ev), _d(vec), ctypes.byref(info)))
        return ev, vec, info

    def bounds(self, steps=60):
        """(Hmin, Hmax) estimated by Lanczos."""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        check(load().evr_plan_bounds(self.h, steps, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    def propagate(self, psi, dt, nsteps, method=CHEBY, npoly=200, poly_tol=1e-12, Hmin=0.0, Hmax=0.0,
                  field=None, mu=None, psi0=None, precision=FP64, autocorr=False, graph=False,
                  field_sub=None):
        """psi(t + nsteps dt) for complex psi (nb_packed,) or (nb_packed, nwp): march_cheby / march_SIL.

        field: (nsteps, nwp) values of the field coupling -E mu (mu a Plan on the
        same basis), constant over each step.  graph: Chebyshev steps after the
        first replayed as one captured HIP graph each (fixed order, exit test
        checked after the replay).  Returns (psi, autocorr or None, info)."""
        r = _to_pairs(psi)
        nwp = r.shape[1] // 2
        fl = None
        if field is not None:
            fl = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(nsteps, nwp))
        p0 = None if psi0 is None else _to_pairs(psi0)
        fs, nsub = None, 0
        if field_sub is not None:       # (nsteps, nsub, nwp): TAYLOR / RK4
            fs = np.ascontiguousarray(field_sub, dtype=np.float64)
            nsub = fs.shape[1]
        o = PropaOpts(method, dt, nsteps, npoly, poly_tol, Hmin, Hmax, _d(fl),
                      None if mu is None else mu.h,
                      None if p0 is None else p0.ctypes.data_as(ctypes.c_void_p), precision, 0,
                      1 if graph else 0, _d(fs), nsub)
        ac = np.zeros((nsteps, nwp, 2)) if autocorr else None
        info = PropaInfo()
        check(load().evr_propagate(self.h, ctypes.byref(o), nwp, r.ctypes.data_as(ctypes.c_void_p), _d(ac),
                                   ctypes.byref(info)))
        out = _from_pairs(r)
        if np.ndim(psi) == 1:
            out = out[:, 0]
        return out, (None if ac is None else ac[..., 0] + 1j * ac[..., 1]), info

    def control_gradient(self, mu, amp, psi0, phi1, T, nsteps, checkpoint=32, npoly=200, poly_tol=1e-14,
                         Hmin=0.0, Hmax=0.0, want=("a",), precision=FP64, graph=False):
        """J_f = |<phi1|psi_f(T)>|^2 for the fields of amp (nk x nf) and its gradients.

        want: any of "a" (nk x nf), "E" (nsteps x nf), "V" (sum_nq x nf).
        precision FP32: sweeps with FP32 H|psi>, J refined in FP64.
        Returns (J, {name: gradient}, info)."""
        amp = np.asfortranarray(np.asarray(amp, dtype=np.float64).reshape(np.shape(amp)[0], -1))
        nk, nf = amp.shape
        o = ControlOpts(T, nsteps, nk, checkpoint, npoly, poly_tol, Hmin, Hmax, precision, 1 if graph else 0)
        p0 = _to_pairs(np.asarray(psi0).reshape(-1, 1))
        p1 = _to_pairs(np.asarray(phi1).reshape(-1, 1))
        J = np.zeros(nf)
        ga = np.zeros((nk, nf), order="F") if "a" in want else None
        gE = np.zeros((nf, nsteps), order="F") if "E" in want else None
        gV = np.zeros((self.basis.sum_nq, nf), order="F") if "V" in want else None
        info = ControlInfo()
        check(load().evr_control_gradient(self.h, mu.h, ctypes.byref(o), nf, _d(amp), _d(p0), _d(p1), _d(J),
                                          _d(ga), _d(gE), _d(gV), ctypes.byref(info)))
        out = {}
        if ga is not None:
            out["a"] = ga
        if gE is not None:
            out["E"] = gE.T.copy()   # (nsteps, nf)
        if gV is not None:
            out["V"] = gV
        return J, out, info

    def BasisRep_TO_GridRep(self, psiB):
        psiB = np.ascontiguousarray(psiB, dtype=np.float64)
        g = np.zeros(self.basis.sum_nq)
        check(load().evr_basis_to_grid(self.h, _d(psiB), _d(g)))
        return g

    def GridRep_TO_BasisRep(self, psiG):
        psiG = np.ascontiguousarray(psiG, dtype=np.float64)
        b = np.zeros(self.nb_packed)
        check(load().evr_grid_to_basis(self.h, _d(psiG), _d(b)))
        return b

    def BasisRep_TO_GridRep_n(self, psiB):
        """BasisTOGrid of the columns of psiB (nb_packed x ncol) -> (sum_nq x ncol)."""
        psiB = np.asfortranarray(psiB, dtype=np.float64)
        if psiB.ndim == 1:
            psiB = psiB.reshape(-1, 1, order="F")
        out = np.zeros((self.basis.sum_nq, psiB.shape[1]), order="F")
        check(load().evr_basis_to_grid_n(self.h, psiB.shape[1], psiB.ctypes.data, out.ctypes.data, 0))
        return out

    def GridRep_TO_BasisRep_n(self, psiG):
        """GridTOBasis of the columns of psiG (sum_nq x ncol) -> (nb_packed x ncol)."""
        psiG = np.asfortranarray(psiG, dtype=np.float64)
        if psiG.ndim == 1:
            psiG = psiG.reshape(-1, 1, order="F")
        out = np.zeros((self.basis.nb_packed, psiG.shape[1]), order="F")
        check(load().evr_grid_to_basis_n(self.h, psiG.shape[1], psiG.ctypes.data, out.ctypes.data, 0))
        return out

    def DerivOp_TO_RVecG(self, psiG, tab_der):
        """Grid derivative of the columns of psiG (DerivOp_TO_RVecG): tab_der =
        (i, j), 1-based modes, 0 for none."""
        psiG = np.asfortranarray(psiG, dtype=np.float64)
        if psiG.ndim == 1:
            psiG = psiG.reshape(-1, 1, order="F")
        out = np.zeros_like(psiG, order="F")
        td = (ctypes.c_int * 2)(*tab_der)
        check(load().evr_grid_deriv(self.h, psiG.shape[1], td, psiG.ctypes.data, out.ctypes.data, 0))
        return out

    def transform_device(self, what, ncol, in_ptr, out_ptr, tab_der=(0, 0)):
        """Device-pointer forms: what = 'b2g', 'g2b' or 'deriv'."""
        L = load()
        if what == "b2g":
            check(L.evr_basis_to_grid_n(self.h, ncol, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr), DEVICE_PTRS))
        elif what == "g2b":
            check(L.evr_grid_to_basis_n(self.h, ncol, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr), DEVICE_PTRS))
        else:
            td = (ctypes.c_int * 2)(*tab_der)
            check(L.evr_grid_deriv(self.h, ncol, td, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr), DEVICE_PTRS))

    def close(self):
        if getattr(self, "h", None):
            load().evr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Multi:
    """Several assembled plans (evr_plan_assemble with ASM_FP64) as one
    block-diagonal operator over the concatenation of their packed vectors
    (evr_multi_create): one H|psi> launch sequence for all systems."""

    def __init__(self, plans, max_nvec=2):
        arr = (ctypes.c_void_p * len(plans))(*[p.h for p in plans])
        h = ctypes.c_void_p()
        check(load().evr_multi_create(arr, len(plans), max_nvec, ctypes.byref(h)))
        self.h = h
        self.n = len(plans)
        off = (ctypes.c_int64 * (self.n + 1))()
        check(load().evr_multi_offsets(self.h, off))
        self.offsets = np.array(off[:], dtype=np.int64)
        self.max_nvec = max_nvec

    @property
    def nb_total(self):
        return int(self.offsets[-1])

    def TabOpPsi(self, psi):
        psi = np.asfortranarray(psi, dtype=np.float64)
        one = psi.ndim == 1
        if one:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        check(load().evr_multi_oppsi(self.h, psi.shape[1], _d(psi), _d(out), 0))
        return out[:, 0] if one else out

    def propagate(self, psi, dt, nsteps, Hmin, Hmax, npoly=200, poly_tol=1e-12):
        """march_cheby for one complex wavepacket per system (psi: the
        concatenation, length nb_total), each system with its own interval."""
        r = np.asfortranarray(np.stack([np.real(psi), np.imag(psi)], axis=1), dtype=np.float64)
        lo = np.ascontiguousarray(Hmin, dtype=np.float64)
        hi = np.ascontiguousarray(Hmax, dtype=np.float64)
        info = PropaInfo()
        check(load().evr_multi_propagate(self.h, dt, nsteps, npoly, poly_tol, _d(lo), _d(hi),
                                         r.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(info)))
        return r[:, 0] + 1j * r[:, 1], info

    def close(self):
        if getattr(self, "h", None):
            load().evr_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Coupled:
    """nb0 coupled electronic states over one basis (evr_coupled_create):
    blocks[i][j] is the plan of H_ij (None for a zero block); a wavepacket is
    nb0 blocks of nb_packed rows, vibrational index fastest."""

    def __init__(self, blocks):
        nb0 = len(blocks)
        self.blocks = [p for row in blocks for p in row]  # keep the plans alive
        arr = (ctypes.c_void_p * (nb0 * nb0))(*[None if p is None else p.h for p in self.blocks])
        h = ctypes.c_void_p()
        check(load().evr_coupled_create(arr, nb0, ctypes.byref(h)))
        self.h = h
        nb = ctypes.c_int64()
        n0 = ctypes.c_int()
        mv = ctypes.c_int()
        check(load().evr_coupled_sizes(self.h, ctypes.byref(nb), ctypes.byref(n0), ctypes.byref(mv)))
        self.nb_packed, self.nb0, self.max_nvec = nb.value, n0.value, mv.value

    def TabOpPsi(self, psi, flags=FP64):
        psi = np.asfortranarray(psi, dtype=np.float64)
        one = psi.ndim == 1
        if one:
            psi = psi.reshape(-1, 1, order="F")
        out = np.zeros_like(psi, order="F")
        for c0 in range(0, psi.shape[1], self.max_nvec):
            c1 = min(psi.shape[1], c0 + self.max_nvec)
            x = np.asfortranarray(psi[:, c0:c1])
            y = np.zeros_like(x, order="F")
            check(load().evr_coupled_oppsi(self.h, c1 - c0, _d(x), _d(y), flags))
            out[:, c0:c1] = y
        return out[:, 0] if one else out

    def diagonal(self):
        """Diagonal of the coupled operator (nb0 * nb_packed,), state-major."""
        d = np.zeros(self.nb0 * self.nb_packed)
        check(load().evr_coupled_diagonal(self.h, _d(d), 0))
        return d

    def davidson(self, nev, block=16, max_sub=0, max_it=200, conv_resi=1e-8, conv_ene=1e-9,
                 precond=True, diag=None, guess=None, want_vectors=False, verbose=False, precision=FP64):
        """Lowest nev vibronic levels by the block Davidson of Plan.davidson
        (FP64 only: the library refuses precision=FP32)."""
        mv = ctypes.c_int()
        check(load().evr_coupled_sizes(self.h, None, None, ctypes.byref(mv)))
        n = self.nb0 * self.nb_packed
        dg = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        gs = None if guess is None else np.asfortranarray(guess, dtype=np.float64)
        o = DavidsonOpts(nev, block, max_sub, max_it, conv_resi, conv_ene, 1 if precond else 0,
                         _d(dg), _d(gs), 0 if gs is None else gs.shape[1], precision, 1e-3,
                         min(block, mv.value), int(verbose))
        ev = np.zeros(nev)
        vec = np.zeros((n, nev), order="F") if want_vectors else None
        info = DavidsonInfo()
        check(load().evr_coupled_davidson(self.h, ctypes.byref(o), _d(ev), _d(vec), ctypes.byref(info)))
        return ev, vec, info

    def propagate(self, psi, dt, nsteps, method=CHEBY, npoly=200, poly_tol=1e-12, Hmin=0.0, Hmax=0.0,
                  psi0=None, autocorr=False, graph=False, field=None, mu=None):
        """march_cheby / march_SIL on the coupled states (see Plan.propagate);
        field (nsteps, nwp) with mu a Coupled dipole operator on the same
        states: step n uses H - field[n] mu."""
        r = _to_pairs(psi)
        nwp = r.shape[1] // 2
        p0 = None if psi0 is None else _to_pairs(psi0)
        fl = None
        if field is not None:
            fl = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(nsteps, nwp))
        o = PropaOpts(method, dt, nsteps, npoly, poly_tol, Hmin, Hmax, _d(fl), None,
                      None if p0 is None else p0.ctypes.data_as(ctypes.c_void_p), FP64, 0, 1 if graph else 0)
        ac = np.zeros((nsteps, nwp, 2)) if autocorr else None
        info = PropaInfo()
        if mu is None:
            check(load().evr_coupled_propagate(self.h, ctypes.byref(o), nwp, r.ctypes.data_as(ctypes.c_void_p),
                                               _d(ac), ctypes.byref(info)))
        else:
            check(load().evr_coupled_propagate_field(self.h, mu.h, ctypes.byref(o), nwp,
                                                     r.ctypes.data_as(ctypes.c_void_p), _d(ac),
                                                     ctypes.byref(info)))
        out = _from_pairs(r)
        if np.ndim(psi) == 1:
            out = out[:, 0]
        return out, (None if ac is None else ac[..., 0] + 1j * ac[..., 1]), info

    def close(self):
        if getattr(self, "h", None):
            load().evr_coupled_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
