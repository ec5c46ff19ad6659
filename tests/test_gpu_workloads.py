"""The control and batch workloads at their full sizes.

- D=6 SG4 L=5 optimal control in atomic units: 32 seeded fields of 64 sine
  coefficients, T = 2000 au in 200 Chebyshev steps of 10 au, checkpoints every
  32 steps.  The adjoint gradients in the coefficients and in the grid values
  of V against central finite differences (h = 1e-6 in a_k, all 32 fields at
  once; h = 1e-6 au in V_g) of the same driver's objective; J against
  |<phi1|psi(T)>|^2 from evr_propagate with the same field; FP32 sweeps with
  the FP64-refined J against FP64 (errors printed).
- 64 seeded systems (D in 6..10, L in 4..6), one wavepacket each, 100
  Chebyshev steps of 10 au as one batched operator: equal to the systems
  propagated one at a time, and the smallest system against the oracle's
  exact propagator exp(-i H T) of the same discrete operator.
"""
import numpy as np
import pytest

import propa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def control_case(evr):
    from evrhip import hamiltonian as H
    D, L = 6, 5
    ham = H.smolyak_config(1, D)
    basis = evr.Basis.sg4(D, L)
    q = basis.grid()
    au = H.AU_TO_CM
    V = ham.potential(q) / au
    plan = evr.Plan(basis, type=10, V=V, G0=ham.G0 / au, Ga=ham.Ga_fortran() / au, max_nvec=64,
                    hermitian=True)
    m1, m2, amp = H.control_config(D)
    mplan = evr.Plan(basis, type=0, V=H.dipole(q, m1, m2), max_nvec=64)
    ev, vec, _ = plan.davidson(2, block=2, conv_resi=1e-9, want_vectors=True)
    plan.assemble(evr.ASM_FP64)
    mplan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    return dict(plan=plan, mplan=mplan, V=V, amp=amp, psi0=vec[:, 0].astype(complex),
                phi1=vec[:, 1].astype(complex), ham=ham, basis=basis)


T, NSTEPS = 2000.0, 200


def _J(c, amp, plan=None):
    p = plan or c["plan"]
    J, _, _ = p.control_gradient(c["mplan"], amp, c["psi0"], c["phi1"], T, NSTEPS, checkpoint=32, want=())
    return J


@pytest.mark.timeout(900)
def test_control_workload_gradients(evr, control_case):
    c = control_case
    amp = c["amp"]
    assert amp.shape == (64, 32)
    J, g, info = c["plan"].control_gradient(c["mplan"], amp, c["psi0"], c["phi1"], T, NSTEPS, checkpoint=32,
                                            want=("a", "V"))
    print("control workload: J mean %.4e, forward %.2f s, backward %.2f s, orders %d" % (
        J.mean(), info.seconds_forward, info.seconds_backward, info.max_order), flush=True)
    # J from the propagation driver with the same mid-step fields
    field = evr.control_field(amp, T, NSTEPS)
    psi = np.repeat(c["psi0"][:, None], amp.shape[1], axis=1)
    out, _, _ = c["plan"].propagate(psi, T / NSTEPS, NSTEPS, npoly=200, poly_tol=1e-14, field=field,
                                    mu=c["mplan"])
    Jp = np.abs(c["phi1"].conj() @ out) ** 2
    assert np.abs(Jp - J).max() < 1e-12
    # finite differences in a_k (all fields at once: they are independent)
    h = 1e-6
    gmax = np.abs(g["a"]).max()
    for k in (0, 7, 31, 63):
        ap, am = amp.copy(), amp.copy()
        ap[k] += h
        am[k] -= h
        fd = (_J(c, ap) - _J(c, am)) / (2 * h)
        err = np.abs(fd - g["a"][k]).max()
        print("  dJ/da_%d: max |fd - adjoint| %.2e (max |grad| %.2e)" % (k, err, gmax), flush=True)
        assert err < 1e-6 * gmax
    # finite differences in V at grid points where the gradient is large
    gV = g["V"]
    pts = np.argsort(-np.abs(gV[:, 0]))[:2].tolist() + np.argsort(-np.abs(gV[:, 1]))[:2].tolist()
    hv = 1e-6
    gvmax = np.abs(gV).max()
    for gp in pts:
        Vp, Vm = c["V"].copy(), c["V"].copy()
        Vp[gp] += hv
        Vm[gp] -= hv
        Js = []
        for Vx in (Vp, Vm):
            c["plan"].set_potential(Vx)          # drops the assembled pools: build them again
            c["plan"].assemble(evr.ASM_FP64)
            Js.append(_J(c, amp))
        fd = (Js[0] - Js[1]) / (2 * hv)
        err = np.abs(fd - gV[gp]).max()
        print("  dJ/dV_%d: max |fd - adjoint| %.2e (max |grad| %.2e)" % (gp, err, gvmax), flush=True)
        assert err < 1e-5 * gvmax
    c["plan"].set_potential(c["V"])
    c["plan"].assemble(evr.ASM_FP64 | evr.ASM_FP32)
    # FP32 sweeps, FP64-refined J
    J32, g32, _ = c["plan"].control_gradient(c["mplan"], amp, c["psi0"], c["phi1"], T, NSTEPS, checkpoint=32,
                                             want=("a",), precision=evr.FP32)
    ea = np.abs(g32["a"] - g["a"]).max() / gmax
    print("FP32-refined vs FP64: max |dJ| %.2e, grad_a rel err %.2e" % (np.abs(J32 - J).max(), ea))
    assert np.abs(J32 - J).max() < 1e-12
    assert ea < 1e-3


def _systems(evr):
    from evrhip import hamiltonian as H
    nsys = 64
    sel = H.SplitMix64(H.config_seed(5), H.FAM_OMEGA).uniform(2 * nsys)
    dims = [6 + min(4, int((u + 1.0) * 2.5)) for u in sel[:nsys]]
    levels = [4 + min(2, int((u + 1.0) * 1.5)) for u in sel[nsys:]]
    return dims, levels


@pytest.mark.timeout(900)
def test_batch_64_systems(evr, oracle_mod):
    from evrhip import hamiltonian as H
    nsteps, dt = 100, 10.0
    dims, levels = _systems(evr)
    assert min(dims) == 6 and max(dims) == 10 and min(levels) == 4 and max(levels) == 6
    plans, lo, hi, hams = [], [], [], []
    for k in range(64):
        ham = H.smolyak_config(100 + k, dims[k]).scaled(1.0 / H.AU_TO_CM)
        b = evr.Basis.sg4(dims[k], levels[k])
        p = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2, hermitian=True)
        p.gen_potential_poly(ham)
        p.assemble(evr.ASM_FP64)
        a, z = p.bounds(60)
        pad = 0.02 * (z - a)
        plans.append(p)
        hams.append(ham)
        lo.append(a - pad)
        hi.append(z + pad)
    m = evr.Multi(plans, max_nvec=2)
    psi = np.concatenate([H.test_vectors(p.nb_packed, 1, 200 + k)[:, 0] for k, p in enumerate(plans)]).astype(complex)
    out, info = m.propagate(psi, dt, nsteps, lo, hi, npoly=200, poly_tol=1e-12)
    print("64 systems, %d packed functions: max order %d" % (m.nb_total, info.max_order), flush=True)
    diff = 0.0
    for k, p in enumerate(plans):
        a, b = m.offsets[k], m.offsets[k + 1]
        ref, _, _ = p.propagate(psi[a:b], dt, nsteps, npoly=200, poly_tol=1e-12, Hmin=lo[k], Hmax=hi[k])
        diff = max(diff, float(np.abs(out[a:b] - ref).max()))
    print("batched vs one by one: %.2e" % diff)
    assert diff < 1e-13
    # the smallest system against exp(-i H T) of the oracle's operator
    k = int(np.argmin([p.nb_packed for p in plans]))
    D, L = dims[k], levels[k]
    s = oracle_mod.SG4(D, L)
    ham = hams[k]
    op = oracle_mod.Op(10, G0=ham.G0, Ga=ham.Ga_fortran(),
                       Vpoly=oracle_mod.poly_coefs(ham.v2, ham.c3, ham.d22, ham.e4))
    M = s.oppsi(op, np.eye(s.nb_packed))
    Hs = 0.5 * (M + M.T)
    a, b = m.offsets[k], m.offsets[k + 1]
    ex = propa.exact_step(Hs, psi[a:b, None], nsteps * dt)[:, 0]
    err = np.abs(out[a:b] - ex).max()
    print("system %d (D=%d L=%d, %d functions) vs exact propagator: %.2e" % (k, D, L, s.nb_packed, err))
    assert err < 1e-9
