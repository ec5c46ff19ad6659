"""CPU restatement of the propagation drivers on a dense operator.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker; the product path never does.

Follows Source_ElVibRot/sub_propagation/sub_module_propa_march.f90 (MARCH):
  cheby_step  march_cheby  MARCH:4142-4433 (E0/Esc MARCH:4211-4231, first
              terms MARCH:4262-4290, recurrence and exit test MARCH:4294-4335,
              Hmax retry MARCH:4348-4373, phase MARCH:4386-4388), with the
              coefficients of cof MARCH:5734-5755 (J_0, 2 J_1, 2 J_2, ...).
  sil_step    march_SIL    MARCH:2899-3100 (three-term Lanczos, E0 = <psi|H|psi>,
              exit on the last Krylov component) with UPsi_spec MARCH:3101-3172.
  exact_step  exp(-i (H - E mu) dt) psi by eigendecomposition: what both
              expansions converge to (the mathematics that pins them).

Pins (tests/test_oracle_propa.py): cheby_step and sil_step against exact_step
on random symmetric matrices and on the symmetrised oracle SG4 operator,
with and without a field; the Chebyshev identity
exp(-i x r) = J_0(r) + 2 sum_k (-i)^k J_k(r) T_k(x) on a grid of x; unitarity.
The Bessel values come from scipy.special.jv (not a restatement).
"""
import numpy as np
from scipy.linalg import expm
from scipy.special import jv


def cheb_coefs(npoly, r):
    """cof(r, npoly): c_0 = J_0(r), c_k = 2 J_k(r), k < npoly."""
    c = jv(np.arange(npoly), r)
    c[1:] *= 2.0
    return c


def _op(H, w, E=None, M=None):
    out = H @ w
    if E is not None:
        out = out - (M @ w) * np.asarray(E)[None, :]
    return out


def cheby_step(H, psi, dt, hmin, hmax, npoly, poly_tol, E=None, M=None, max_cheby=10):
    """One march_cheby step for psi (nb, nwp) complex.  Returns (psi, order, hmax)."""
    psi = np.asarray(psi, dtype=np.complex128)
    for _ in range(max_cheby):
        dE = hmax - hmin
        E0 = hmin + 0.5 * dE
        esc = 0.5 * dE
        c = cheb_coefs(npoly, esc * dt)
        w1 = psi.copy()
        out = c[0] * psi
        w2 = -1j * (_op(H, w1, E, M) - E0 * w1) / esc
        out = out + c[1] * w2
        ok = False
        k = 1
        for k in range(2, npoly):
            w3 = w1 - 2j * (_op(H, w2, E, M) - E0 * w2) / esc
            w1, w2 = w2, w3
            out = out + c[k] * w2
            crit = np.max(np.abs((np.abs(w2) ** 2).sum(0) * c[k]))
            if crit > 1e15:
                break
            if crit < poly_tol:
                ok = True
                break
        if ok:
            return out * np.exp(-1j * E0 * dt), k, hmax
        hmax = hmax + 0.1 * dE
    raise RuntimeError("march_cheby did not converge")


def _upsi_spec(T, dt):
    ev, V = np.linalg.eigh(T)
    return V @ (np.exp(-1j * ev * dt) * V[0, :])


def sil_step(H, psi, dt, npoly, poly_tol, E=None, M=None):
    """One march_SIL step (reference recurrence: no re-orthogonalisation)."""
    psi = np.asarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    orders = []
    for w in range(psi.shape[1]):
        Ew = None if E is None else [E[w]]
        nu = np.linalg.norm(psi[:, w])
        q = [psi[:, w] / nu]
        T = np.zeros((npoly, npoly))
        E0 = 0.0
        n = 0
        u = None
        for k in range(npoly):
            w1 = _op(H, q[k][:, None], Ew, M)[:, 0]
            if k == 0:
                E0 = np.vdot(q[0], w1).real
            w1 = w1 - E0 * q[k]
            if k > 0:
                w1 = w1 - T[k, k - 1] * q[k - 1]
            a = np.vdot(w1, q[k])
            T[k, k] = a.real
            u = _upsi_spec(T[:k + 1, :k + 1], dt)
            if abs(u[k]) < poly_tol or k == npoly - 1:
                n = k + 1
                break
            w1 = w1 - a * q[k]
            b = np.linalg.norm(w1)
            T[k + 1, k] = T[k, k + 1] = b
            q.append(w1 / b)
        out[:, w] = nu * np.exp(-1j * E0 * dt) * sum(u[k] * q[k] for k in range(n))
        orders.append(n)
    return out, max(orders)


def exact_step(H, psi, dt, E=None, M=None):
    """exp(-i (H - E_w M) dt) psi_w by eigendecomposition."""
    psi = np.asarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    if E is None:
        ev, V = np.linalg.eigh(H)
        return V @ (np.exp(-1j * ev * dt)[:, None] * (V.T @ psi))
    for w in range(psi.shape[1]):
        ev, V = np.linalg.eigh(H - E[w] * M)
        out[:, w] = V @ (np.exp(-1j * ev * dt) * (V.T @ psi[:, w]))
    return out


def taylor_field_step(H, psi, dt, E0, npoly, poly_tol, dnE=None, M=None):
    """One march_noD_field step (MARCH:1635-1836) for psi (nb, nwp): the
    Taylor series of i dpsi/dt = (H - E0 - E(t) M) psi about t_n with the
    field derivatives dnE[k, w] = E^(k)(t_n):
      w_0 = psi, w_{i+1} = (-i dt/(i+1)) [(H - E0) w_i
                 + M sum_{k <= min(i, max_der)} (dt^k/k!) (-E^(k)) w_{i-k}],
    max_der the largest k with max|E^(k)| dt^(k+1)/k! > poly_tol (MARCH:1698-1711),
    stopping once |w_{i+1}|^2 < poly_tol for every column (this restatement
    stops when all columns have; the reference loops per wavepacket);
    psi(t_n + dt) = e^{-i E0 dt} sum_i w_i.  Returns (psi, order)."""
    psi = np.asarray(psi, dtype=np.complex128)
    fk = None if dnE is None else -np.asarray(dnE, dtype=np.float64)
    max_der = 0
    if fk is not None:
        fac = dt
        for k in range(1, min(len(fk), npoly)):
            fac *= dt / k
            if np.abs(fk[k]).max() * fac > poly_tol:
                max_der = k
    work = [psi.copy()]
    out = psi.copy()
    order = 0
    for i in range(npoly - 1):
        hw = H @ work[i] - E0 * work[i]
        if fk is not None and np.abs(fk).max() >= poly_tol:
            s = np.zeros_like(psi)
            f = 1.0
            for k in range(0, min(i, max_der) + 1):
                if k > 0:
                    f *= dt / k
                s = s + f * fk[k][None, :] * work[i - k]
            hw = hw + M @ s
        wn = (-1j * dt / (i + 1)) * hw
        work.append(wn)
        out = out + wn
        order = i + 1
        if np.all((np.abs(wn) ** 2).sum(0) < poly_tol):
            break
    return out * np.exp(-1j * E0 * dt), order


def rk4_field_step(H, psi, dt, E0, Evals=None, M=None):
    """One march_RK4_field step (MARCH:392-543) with fcn_field
    (MARCH:1434-1515): f(t, x) = -i [(H - E0) x - E(t) M x], classical RK4
    with the field at t_n, t_n + dt/2, t_n + dt (Evals[j, w]), then the phase
    e^{-i E0 dt}."""
    psi = np.asarray(psi, dtype=np.complex128)

    def f(x, j):
        y = H @ x - E0 * x
        if Evals is not None:
            y = y - (M @ x) * np.asarray(Evals[j])[None, :]
        return -1j * y
    w1 = f(psi, 0)
    w3 = f(psi + w1 * dt / 2, 1)
    w4 = f(psi + w3 * dt / 2, 1)
    w2 = psi + w4 * dt
    w4 = w3 + w4
    w3 = f(w2, 2)
    return (psi + (w1 + w3 + 2 * w4) * dt / 6) * np.exp(-1j * E0 * dt)


def magnus4_propagate(H, psi, t0, t1, nsub, field, M):
    """psi(t1) of i dpsi/dt = (H - E(t) M) psi by nsub fourth-order Magnus
    steps (two Gauss points): the reference solution of the time-dependent
    equation that the field integrators converge to."""
    psi = np.asarray(psi, dtype=np.complex128)
    h = (t1 - t0) / nsub
    c = np.sqrt(3.0) / 6.0
    for n in range(nsub):
        ta, tb = t0 + (n + 0.5 - c) * h, t0 + (n + 0.5 + c) * h
        Ha = H - field(ta) * M
        Hb = H - field(tb) * M
        Om = -1j * h / 2 * (Ha + Hb) - (np.sqrt(3.0) / 12.0) * h * h * (Hb @ Ha - Ha @ Hb)
        psi = expm(Om) @ psi
    return psi
