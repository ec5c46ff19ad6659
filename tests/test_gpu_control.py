"""Control objective and adjoint gradient on the GPU against the oracle.

System: D=3 SG4 L=3 (type 10, symmetric form) in atomic units, a dipole
surface mu(q) on the grid as a type-0 plan, fields
E(t) = sin^2(pi t/T) sum_k a_k sin(k pi t/T) held at mid-step values.
The oracle propagates with exact exp(-i (H - E mu) dt) and differentiates
with the Daleckii-Krein formula; the GPU differentiates its own Chebyshev
recurrences.  Bars: gradients within 1e-9 of the oracle (relative to the
largest entry), central differences (h = 1e-6) of the GPU objective within
1e-6, identical results for any checkpoint interval.
"""
import numpy as np
import pytest

import control

pytestmark = pytest.mark.gpu

T, N, NK = 400.0, 20, 4


@pytest.fixture(scope="module")
def case(evr, oracle_mod):
    from evrhip import hamiltonian as Hm
    D, L = 3, 3
    ham = Hm.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    V = ham.potential(q) / Hm.AU_TO_CM
    G0 = ham.G0 / Hm.AU_TO_CM
    Ga = ham.Ga_fortran() / Hm.AU_TO_CM
    M = s.oppsi(oracle_mod.Op(10, V=V, G0=G0, Ga=Ga), np.eye(s.nb_packed))
    H = 0.5 * (M + M.T)
    mu = 0.1 * q[:, 0] + 0.05 * q[:, 1] - 0.03 * q[:, 0] * q[:, 2]
    Mu = s.oppsi(oracle_mod.Op(0, V=mu), np.eye(s.nb_packed))
    Mu = 0.5 * (Mu + Mu.T)
    ev, Vec = np.linalg.eigh(H)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, V=V, G0=G0, Ga=Ga, max_nvec=8, hermitian=True)
    mplan = evr.Plan(b, type=0, V=mu, max_nvec=8)
    rng = np.random.default_rng(11)
    amp = rng.uniform(-0.02, 0.02, (NK, 3))
    return dict(s=s, H=H, Mu=Mu, plan=plan, mplan=mplan, psi0=Vec[:, 0].astype(complex),
                phi1=Vec[:, 1].astype(complex), amp=amp, ev=ev)


def run(c, amp, want=("a", "E", "V"), checkpoint=8):
    return c["plan"].control_gradient(c["mplan"], amp, c["psi0"], c["phi1"], T, N, checkpoint=checkpoint,
                                      npoly=100, poly_tol=1e-15, want=want)


def test_field_values(evr, case):
    assert np.abs(evr.control_field(case["amp"], T, N) - control.field_values(case["amp"], T, N)).max() < 1e-16


def test_gradient_vs_oracle(evr, case):
    c = case
    J, g, info = run(c, c["amp"])
    E = control.field_values(c["amp"], T, N)
    Y = control.grid_basis(c["s"])
    gw = control.grid_weights(c["s"])
    Jr, gEr, gVr = control.gradient(c["H"], c["Mu"], c["psi0"], c["phi1"], E, T / N, Y, gw)
    gar = control.field_basis(T, N, NK).T @ gEr
    print("J", J, "orders", info.max_order, "hpsi", info.hpsi)
    assert np.abs(J - Jr).max() < 1e-12
    assert np.abs(g["E"] - gEr).max() < 1e-9 * np.abs(gEr).max()
    assert np.abs(g["a"] - gar).max() < 1e-9 * np.abs(gar).max()
    assert np.abs(g["V"] - gVr).max() < 1e-9 * np.abs(gVr).max()


def test_gradient_finite_differences(evr, case):
    c = case
    _, g, _ = run(c, c["amp"], want=("a",))
    h = 1e-6
    for k in [0, 2]:
        ap, am = c["amp"].copy(), c["amp"].copy()
        ap[k] += h
        am[k] -= h
        Jp, _, _ = run(c, ap, want=())
        Jm, _, _ = run(c, am, want=())
        fd = (Jp - Jm) / (2 * h)
        assert np.abs(fd - g["a"][k]).max() < 1e-6 * np.abs(g["a"]).max()


def test_checkpoint_interval_invariance(evr, case):
    c = case
    _, g1, _ = run(c, c["amp"], want=("E", "V"), checkpoint=1)
    _, g2, _ = run(c, c["amp"], want=("E", "V"), checkpoint=32)
    assert np.abs(g1["E"] - g2["E"]).max() <= 1e-14 * np.abs(g1["E"]).max()
    assert np.abs(g1["V"] - g2["V"]).max() <= 1e-14 * np.abs(g1["V"]).max()


def test_objective_matches_propagation(evr, case):
    """J from the control driver equals |<phi1|psi(T)>|^2 of evr_propagate with the same field."""
    c = case
    E = evr.control_field(c["amp"], T, N)
    J, _, _ = run(c, c["amp"], want=())
    psi = np.repeat(c["psi0"][:, None], E.shape[1], axis=1)
    out, _, _ = c["plan"].propagate(psi, T / N, N, npoly=100, poly_tol=1e-15, field=E, mu=c["mplan"])
    assert np.abs(np.linalg.norm(out, axis=0) - 1.0).max() < 1e-12      # unitarity
    assert np.abs(np.abs(c["phi1"].conj() @ out) ** 2 - J).max() < 1e-10


def test_autograd_function_wrapper(evr, case):
    """torch.autograd.Function over the C ABI (test harness only), checked by gradcheck."""
    import torch
    c = case

    class ControlJ(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            J, g, _ = run(c, a.detach().numpy(), want=("a",))
            ctx.save_for_backward(torch.from_numpy(g["a"]))
            return torch.from_numpy(J)

        @staticmethod
        def backward(ctx, gJ):
            (ga,) = ctx.saved_tensors
            return ga * gJ[None, :]

    a = torch.tensor(c["amp"][:, :1], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(ControlJ.apply, (a,), eps=1e-6, atol=1e-8, rtol=1e-5)


def test_fp32_refined_gradient(evr, case):
    """precision FP32: both sweeps apply H and mu in FP32, J is refined by an
    FP64 forward sweep (equal to the FP64 run's J to 1e-12); the gradients
    carry the FP32 error of H|psi> (1e-4 relative here)."""
    c = case
    J64, g64, _ = run(c, c["amp"], want=("a", "E"))
    J32, g32, info = c["plan"].control_gradient(c["mplan"], c["amp"], c["psi0"], c["phi1"], T, N, checkpoint=8,
                                                npoly=100, poly_tol=1e-15, want=("a", "E"),
                                                precision=evr.FP32)
    ea = np.abs(g32["a"] - g64["a"]).max() / np.abs(g64["a"]).max()
    print("FP32-refined control: |dJ| %.2e, grad_a rel err %.2e" % (np.abs(J32 - J64).max(), ea))
    assert np.abs(J32 - J64).max() < 1e-12
    assert ea < 1e-4


def test_graph_forward_sweep_identical(evr, case):
    """Forward steps replayed as a captured Chebyshev step (graph = 1): J and
    the gradients equal the eager sweep's to the last bit."""
    c = case
    J0, g0, _ = run(c, c["amp"], want=("a", "E"))
    J1, g1, _ = c["plan"].control_gradient(c["mplan"], c["amp"], c["psi0"], c["phi1"], T, N, checkpoint=8,
                                           npoly=100, poly_tol=1e-15, want=("a", "E"), graph=True)
    assert np.array_equal(J0, J1)
    assert np.array_equal(g0["a"], g1["a"]) and np.array_equal(g0["E"], g1["E"])


def test_build_field_matches_reference_update(evr, case):
    """build_field from the GPU adjoint gradient, (dJ/dE_n)/(2 alpha dt)
    s(t_n) (the reference formula -(1/alpha) s Im(<psi|chi><chi|mu|psi>), to
    which the oracle ties it), against the same from the oracle's exact
    gradient (1e-9); a small step along it raises J; the Krotov form adds the
    old field."""
    c = case
    N2 = 100
    amp = c["amp"][:, :1]
    J, g, _ = c["plan"].control_gradient(c["mplan"], amp, c["psi0"], c["phi1"], T, N2, checkpoint=20,
                                         npoly=100, poly_tol=1e-15, want=("E",))
    alpha = 0.5
    Enew = evr.control_build_field(g["E"], T, alpha, Tenv=T)
    E = evr.control_field(amp, T, N2)
    _, gEr, _ = control.gradient(c["H"], c["Mu"], c["psi0"], c["phi1"], E, T / N2)
    tn = (np.arange(N2) + 0.5) * (T / N2)
    ref = gEr[:, 0] / (2 * alpha * T / N2) * np.sin(np.pi * tn / T) ** 2
    err = np.abs(Enew[:, 0] - ref).max() / np.abs(ref).max()
    print("build_field vs oracle: rel err %.2e" % err)
    assert err < 1e-9
    kro = evr.control_build_field(g["E"], T, alpha, Tenv=T, field_old=E)
    assert np.array_equal(kro, E + Enew)
    # ascent: J(E + eps E_new) > J(E) for a small eps
    psi = c["psi0"][:, None]
    J0 = np.abs(c["phi1"].conj() @ c["plan"].propagate(psi, T / N2, N2, npoly=100, poly_tol=1e-15, field=E,
                                                       mu=c["mplan"])[0]) ** 2
    eps = 1e-3 * np.abs(E).max() / np.abs(Enew).max()
    J1 = np.abs(c["phi1"].conj() @ c["plan"].propagate(psi, T / N2, N2, npoly=100, poly_tol=1e-15,
                                                       field=E + eps * Enew, mu=c["mplan"])[0]) ** 2
    assert J1[0] > J0[0]
