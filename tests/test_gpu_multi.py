"""Several systems of different dimension and level as one block-diagonal
operator (evr_multi_*): H|psi> and Chebyshev propagation against the same
plans run one by one (the single-system paths are pinned to the oracle
elsewhere), and one system against the exact propagator of the oracle's
operator.  Bars: H|psi> 1e-14 relative, psi(t) 1e-12 against the one-by-one
runs and 1e-10 against the exact propagation.
"""
import numpy as np
import pytest

import propa

pytestmark = pytest.mark.gpu

SYSTEMS = [(3, 4), (4, 3), (5, 3), (2, 6), (4, 4)]


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def systems(evr):
    from evrhip import hamiltonian as H
    out = []
    for k, (D, L) in enumerate(SYSTEMS):
        ham = H.smolyak_config(20 + k, D)
        b = evr.Basis.sg4(D, L)
        plan = evr.Plan(b, type=10, V=ham.potential(b.grid()), G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2,
                        hermitian=True)
        plan.assemble(evr.ASM_FP64)
        lo, hi = plan.bounds(60)
        pad = 0.02 * (hi - lo)
        out.append(dict(ham=ham, basis=b, plan=plan, lo=lo - pad, hi=hi + pad))
    return out


def test_multi_oppsi(evr, systems):
    m = evr.Multi([s["plan"] for s in systems], max_nvec=2)
    assert m.n == len(systems)
    X = evr.hamiltonian.test_vectors(m.nb_total, 2, 11)
    got = m.TabOpPsi(X)
    for k, s in enumerate(systems):
        a, b = m.offsets[k], m.offsets[k + 1]
        ref = s["plan"].TabOpPsi(np.asfortranarray(X[a:b]))
        assert relerr(got[a:b], ref) < 1e-14


def test_multi_propagate(evr, oracle_mod, systems):
    m = evr.Multi([s["plan"] for s in systems], max_nvec=2)
    rng = np.random.default_rng(12)
    psi = rng.standard_normal(m.nb_total) + 1j * rng.standard_normal(m.nb_total)
    for k in range(m.n):
        a, b = m.offsets[k], m.offsets[k + 1]
        psi[a:b] /= np.linalg.norm(psi[a:b])
    lo = [s["lo"] for s in systems]
    hi = [s["hi"] for s in systems]
    dt, nst = 2e-3, 6
    out, info = m.propagate(psi, dt, nst, lo, hi, npoly=300, poly_tol=1e-14)
    orders = []
    for k, s in enumerate(systems):
        a, b = m.offsets[k], m.offsets[k + 1]
        ref, _, inf = s["plan"].propagate(psi[a:b], dt, nst, npoly=300, poly_tol=1e-14, Hmin=lo[k], Hmax=hi[k])
        orders.append(inf.max_order)
        print("system", SYSTEMS[k], "order", inf.max_order, "diff", np.abs(out[a:b] - ref).max())
        assert np.abs(out[a:b] - ref).max() < 1e-12
    assert info.max_order == max(orders)
    # one system against the exact propagator of the oracle's operator
    D, L = SYSTEMS[1]
    s = oracle_mod.SG4(D, L)
    ham = systems[1]["ham"]
    M = s.oppsi(oracle_mod.Op(10, V=ham.potential(s.grid()), G0=ham.G0, Ga=ham.Ga_fortran()),
                np.eye(s.nb_packed))
    ex = psi[m.offsets[1]:m.offsets[2]][:, None]
    for _ in range(nst):
        ex = propa.exact_step(0.5 * (M + M.T), ex, dt)
    assert np.abs(out[m.offsets[1]:m.offsets[2]] - ex[:, 0]).max() < 1e-10
