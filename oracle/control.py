"""CPU restatement of the control objective and its exact gradient (dense).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker; the product path never does.

The objective J_f = |<phi1|psi_f(T)>|^2 is the fidelity of calc_fidelity
(Source_ElVibRot/sub_propagation/sub_control.f90:1000-1057); the field form
E(t) = sin^2(pi t/T) sum_k a_k sin(k pi t/T), held at its mid-step value over
each step, and the gradient are not in the reference.  Here every step is the
exact propagator U_n = exp(-i (H - E_n mu) dt) from an eigendecomposition and
dU_n/dtheta comes from the Daleckii-Krein formula
    dU = V [F o (V^T dA V)] V^T,  F_ij = (f(l_i) - f(l_j)) / (l_i - l_j),
    F_ii = f'(l_i),  f(l) = exp(-i l dt),
with dA = -mu for a field value and dA = P_g = gw_g y_g y_g^T for a grid value
of V (y_g: the basis functions at grid point g, gw_g: Smolyak weight times
quadrature weight).  Pins (tests/test_oracle_control.py): central finite
differences of J in E_n and in V_g, and P_g against the oracle's own type-0
operator with V = e_g.
"""
import numpy as np


def field_basis(T, nsteps, nk):
    """S[n, k-1] = sin^2(pi t_n/T) sin(k pi t_n/T), t_n = (n + 1/2) T/nsteps."""
    t = (np.arange(nsteps) + 0.5) * (T / nsteps)
    k = np.arange(1, nk + 1)
    return (np.sin(np.pi * t / T) ** 2)[:, None] * np.sin(np.outer(t, k) * np.pi / T)


def field_values(amp, T, nsteps):
    amp = np.asarray(amp, dtype=np.float64).reshape(np.shape(amp)[0], -1)
    return field_basis(T, nsteps, amp.shape[0]) @ amp


def field_derivs(amp, T, t, nder):
    """E^(j)(t), j = 0..nder, of E(t) = sin^2(pi t/T) sum_k a_k sin(k pi t/T)
    (amp: nk or nk x nf) from sin^2(x) sin(kx) = sin(kx)/2 - sin((k+2)x)/4
    - sin((k-2)x)/4 and d^j/dt^j sin(m w t) = (m w)^j sin(m w t + j pi/2),
    w = pi/T.  Returns (nder + 1, nf)."""
    amp = np.asarray(amp, dtype=np.float64).reshape(np.shape(amp)[0], -1)
    nk = amp.shape[0]
    w = np.pi / T
    out = np.zeros((nder + 1, amp.shape[1]))
    for j in range(nder + 1):
        for k in range(1, nk + 1):
            v = 0.0
            for m, c in ((k, 0.5), (k + 2, -0.25), (k - 2, -0.25)):
                v += c * (m * w) ** j * np.sin(m * w * t + j * np.pi / 2)
            out[j] += v * amp[k - 1]
    return out


def grid_weights(s):
    """gw (sum_nq) of an oracle SG4 basis: term weight x tensor quadrature weight
    (W[0, q] / B[q, 0] per 1D rule, first mode fastest inside a term)."""
    lv, w, _, onq = s.terms()
    gw = np.zeros(s.sum_nq)
    cache = {}
    for t in range(s.nterms):
        g = np.array([w[t]])
        for d in range(s.D):
            key = (d, int(lv[t, d]))
            if key not in cache:
                o = s.one_d(d, int(lv[t, d]))
                cache[key] = o["W"][0, :] / o["B"][:, 0]
            g = np.kron(cache[key], g)
        gw[onq[t]:onq[t + 1]] = g
    return gw


def grid_basis(s):
    """Y (sum_nq x nb_packed): every packed basis function on the term grids."""
    return np.stack([s.b2g(e) for e in np.eye(s.nb_packed)], axis=1)


def _dk(ev, dt):
    f = np.exp(-1j * ev * dt)
    d = ev[:, None] - ev[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        F = (f[:, None] - f[None, :]) / d
    small = np.abs(d) < 1e-9 * (1.0 + np.abs(ev).max())
    fp = -1j * dt * f
    F[small] = (0.5 * (fp[:, None] + fp[None, :]))[small]
    return f, F


def objective(H, Mu, psi0, phi1, field, dt):
    """J_f for every column f of field (nsteps, nf)."""
    nsteps, nf = field.shape
    J = np.zeros(nf)
    for f in range(nf):
        psi = np.asarray(psi0, dtype=np.complex128).copy()
        for n in range(nsteps):
            ev, V = np.linalg.eigh(H - field[n, f] * Mu)
            psi = V @ (np.exp(-1j * ev * dt) * (V.T @ psi))
        J[f] = abs(np.vdot(phi1, psi)) ** 2
    return J


def gradient(H, Mu, psi0, phi1, field, dt, Y=None, gw=None):
    """(J, dJ/dE (nsteps, nf), dJ/dV (sum_nq, nf) or None)."""
    nsteps, nf = field.shape
    J = np.zeros(nf)
    gE = np.zeros((nsteps, nf))
    gV = None if Y is None else np.zeros((Y.shape[0], nf))
    for f in range(nf):
        eig = []
        psis = [np.asarray(psi0, dtype=np.complex128).copy()]
        for n in range(nsteps):
            ev, V = np.linalg.eigh(H - field[n, f] * Mu)
            eig.append((ev, V))
            psis.append(V @ (np.exp(-1j * ev * dt) * (V.T @ psis[-1])))
        o = np.vdot(phi1, psis[-1])
        J[f] = abs(o) ** 2
        g = 2.0 * o * np.asarray(phi1, dtype=np.complex128)   # dJ = Re<g, dpsi(T)>
        for n in range(nsteps - 1, -1, -1):
            ev, V = eig[n]
            fv, F = _dk(ev, dt)
            gt = V.T @ g
            pt = V.T @ psis[n]
            X = np.conj(gt)[:, None] * F * pt[None, :]        # Re sum_ij X_ij (V^T dA V)_ij
            gE[n, f] = np.real(np.sum(X * (V.T @ (-Mu) @ V)))
            if Y is not None:
                YV = Y @ V
                gV[:, f] += gw * np.real(np.sum((YV @ X) * YV, axis=1))
            g = V @ (np.conj(fv) * gt)                           # U_n^dagger g
    return J, gE, gV


def build_field(H, Mu, psi0, phi1, field, dt, alpha, Tenv=0.0, field_old=None):
    """build_field (CTRL:902-971) for one field held at mid-step values:
    E_new(t_n) = [E_old] - (1/alpha) s(t_n) Im(<psi|chi> <chi|mu|psi>) at the
    mid-steps t_n = (n + 1/2) dt, psi propagated forward from psi0 and chi
    backward from phi1 exactly (exp(-i (H - E mu) dt) per step, half steps to
    reach the middle), s = sin^2(pi t/Tenv) (envelopp) or 1."""
    field = np.asarray(field, dtype=np.float64).ravel()
    N = len(field)

    def U(E, tau):
        ev, V = np.linalg.eigh(H - E * Mu)
        return (V * np.exp(-1j * ev * tau)) @ V.conj().T
    psi = [np.asarray(psi0, complex)]
    for n in range(N):
        psi.append(U(field[n], dt) @ psi[-1])
    chi = [None] * (N + 1)
    chi[N] = np.asarray(phi1, complex)
    for n in range(N - 1, -1, -1):
        chi[n] = U(field[n], -dt) @ chi[n + 1]
    out = np.zeros(N)
    for n in range(N):
        ph = U(field[n], dt / 2) @ psi[n]
        ch = U(field[n], -dt / 2) @ chi[n + 1]
        S = np.vdot(ph, ch)
        avmu = np.vdot(ch, Mu @ ph)
        t = (n + 0.5) * dt
        env = np.sin(np.pi * t / Tenv) ** 2 if Tenv > 0 else 1.0
        out[n] = -np.imag(S * avmu) / alpha * env + (0.0 if field_old is None else field_old[n])
    return out
