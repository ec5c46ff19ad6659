#!/usr/bin/env python3
"""H|psi> throughput of the MI355X engine on the D=9 Smolyak SG4 L=8 workload.

One step = one H|psi> application (type-10 Hamiltonian, q-linear G) to a block
of `--nvec` vectors resident in HBM, through the C ABI (evr_oppsi with device
pointers).  The value is H|psi> applications per second per vector
(vectors x steps / time).  Prints one JSON line.

The CPU baseline leg times the oracle (a C/OpenMP restatement of the
reference's per-term OpenMP loop, PSG4_omp=1) on a strided sample of the same
terms on the host cores and scales by the grid-point fraction of the sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "elvibrot-tnumtana_amd"))

PEAK_FP64 = 78.6   # TFLOP/s, MI355X dense FP64 (matrix and vector are equal on gfx950)
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32
PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E


def config_index(D, L):
    """Seed family of the Hamiltonian generator: the D=6 L=5, D=9 L=8 and
    D=12 L=6 Smolyak workloads have their own; any other size uses family 9."""
    return {(6, 5): 1, (9, 8): 2, (12, 6): 4}.get((D, L), 9)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model, logical CPUs of the host, physical cores of the host, and
    the CPUs this process may run on (its share of the host)."""
    model, phys = "unknown", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k in ("physical id", "core id"):
                cur[k] = v
            elif not k and cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
    except OSError:
        pass
    share = len(os.sched_getaffinity(0))
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "cpus_allowed": share}


def cpu_leg(args):
    """(Child process.)  The oracle's H|psi> on one vector over EVERY Smolyak
    term of the workload, timed on the host cores: the reference's per-term
    OpenMP loop (PSG4_omp=1) restated in C, built -O3 -march=native for this
    host.  Prints one JSON object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from evrhip import hamiltonian as H
    D, L = args.D, args.L
    ham = H.smolyak_config(config_index(D, L), D)
    s = oracle.SG4(D, L)
    # the reference keeps V on the grid (Save_MemGrid) and reads it per call;
    # a zero-filled (lazily mapped) array gives the same loads and flops
    V = np.zeros(s.sum_nq)
    op = oracle.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran())
    X = H.test_vectors(s.nb_packed, 1, 2)
    threads = int(os.environ["OMP_NUM_THREADS"])
    s.oppsi(op, X, nthreads=threads, t0=0, t1=min(s.nterms, 64))     # warm-up (pages, threads)
    t0 = time.perf_counter()
    s.oppsi(op, X, nthreads=threads)
    dt = time.perf_counter() - t0
    print(json.dumps({"seconds": dt, "threads": threads}), flush=True)


def cpu_baseline(D, L, threads):
    """Builds the oracle with -O3 -march=native for this host (a private
    copy under /tmp) and times its full H|psi> in a child process with
    OMP_PROC_BIND=close, OMP_PLACES=cores."""
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="evr_cpu_")
    lib = os.path.join(tmp, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-std=c11", "-shared", "-o", lib,
                    os.path.join(ROOT, "oracle", "evr_oracle.c"), "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
               ORACLE_LIB=lib)
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", "--D", str(D), "--L", str(L)],
                         env=env, check=True, capture_output=True, text=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    hc = host_cpu()
    return {"value": 1.0 / r["seconds"], "unit": "Hpsi/s per vector", "cores": r["threads"], "kind": "port",
            "sample": "oracle (C/OpenMP restatement of the reference per-term loop with atomic scatter, "
                      "FP64, gcc -O3 -march=native), one vector, every Smolyak term (full H|psi>), "
                      "%.1f s wall, OMP_PROC_BIND=close OMP_PLACES=cores" % r["seconds"],
            "host": hc}


def copy_calibration():
    """1 GiB device-to-device copies: their FETCH_SIZE / WRITE_SIZE per dispatch
    give the KiB-to-bytes scale of the HBM counters for a known traffic."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}), flush=True)


def davidson_pair(plan, nev, nvec, conv_resi, refine_resi, max_it, verbose=True):
    """The nev lowest levels of a symmetric plan by block Davidson, plain FP64
    and mixed (FP32 H|psi> of the new vectors, FP64 defect corrections):
    per run the iterations, H|psi> counts and wall, and max |dE| between them."""
    import numpy as np
    import evrhip
    out = {}
    for tag, prec in (("fp64", evrhip.FP64), ("fp32_refined", evrhip.FP32)):
        t0 = time.perf_counter()
        ev, _, info = plan.davidson(nev, block=nvec, conv_resi=conv_resi, conv_ene=1e-7, max_it=max_it,
                                    precision=prec, refine_resi=refine_resi, verbose=verbose)
        wall = time.perf_counter() - t0
        out[tag] = dict(ev=ev, it=info.iterations, conv=info.converged, h64=info.hpsi_fp64,
                        h32=info.hpsi_fp32, resi=info.max_resi, s=wall)
        log("%s: %d iterations, converged=%d, H|psi> fp64 %d fp32 %d, max residual %.2e, %.2f s"
            % (tag, info.iterations, info.converged, info.hpsi_fp64, info.hpsi_fp32, info.max_resi, wall))
    dE = np.abs(out["fp64"]["ev"] - out["fp32_refined"]["ev"])
    res = {"max_abs_dE_cm-1": float(dE.max()), "mean_abs_dE_cm-1": float(dE.mean()),
           "E0_cm-1": float(out["fp64"]["ev"][0]), "E_last_cm-1": float(out["fp64"]["ev"][-1]),
           "nev": nev, "block": nvec, "conv_resi_cm-1": conv_resi,
           "dE_per_level_cm-1": [float("%.3g" % x) for x in dE]}
    for tag in ("fp64", "fp32_refined"):
        o = out[tag]
        res[tag] = {"iterations": o["it"], "converged": o["conv"], "hpsi_fp64": o["h64"],
                    "hpsi_fp32": o["h32"], "fp32_share": round(o["h32"] / max(1, o["h32"] + o["h64"]), 4),
                    "max_residual": o["resi"], "seconds": round(o["s"], 3)}
    res["wall_speedup"] = round(out["fp64"]["s"] / out["fp32_refined"]["s"], 3)
    return res


def davidson_levels(args):
    """The lowest --nev levels by block Davidson on the symmetric form of the
    workload operator, plain FP64 and mixed FP32/FP64: max |dE| between them."""
    import evrhip
    from evrhip import hamiltonian as H
    D, L = args.D, args.L
    cfg_index = config_index(D, L)
    ham = H.smolyak_config(cfg_index, D)
    basis = evrhip.Basis.sg4(D, L)
    plan = evrhip.Plan(basis, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=args.nvec, hermitian=True)
    plan.gen_potential_poly(ham)
    t_asm = None
    if not args.no_assembled:
        try:
            t0 = time.perf_counter()
            plan.assemble(evrhip.ASM_FP64 | evrhip.ASM_FP32)
            t_asm = time.perf_counter() - t0
            log("assembled operator (FP64 + FP32) in %.1f s" % t_asm)
        except RuntimeError as e:
            log("assembled operator skipped: %s" % e)
    line = {"workload": "SG4 D=%d L=%d symmetric type-10, %d lowest levels, block %d" % (D, L, args.nev, args.nvec),
            "operator": "sum-factorised term kernels" if t_asm is None else
                        "assembled per-term operators (built in %.1f s, not in the Davidson times)" % t_asm}
    line.update(davidson_pair(plan, args.nev, args.nvec, args.conv_resi, args.refine_resi, args.max_it))
    print(json.dumps(line), flush=True)


def control_run(args):
    """Optimal control on D=6 SG4 L=5 in atomic units: 32 seeded fields of 64
    sine coefficients, T = 2000 au in 200 Chebyshev steps of 10 au, objective
    |<phi_1|psi_f(T)>|^2 from the ground state, gradient in the coefficients
    and the V grid values by the adjoint sweep with checkpoints every 32 steps;
    then the 32 forward propagations alone, adaptive and as captured graphs.
    Sum-factorised operators first, then the assembled ones."""
    import numpy as np
    import evrhip
    from evrhip import hamiltonian as H
    D, L, T, nsteps = 6, 5, 2000.0, 200
    ham = H.smolyak_config(1, D)
    basis = evrhip.Basis.sg4(D, L)
    q = basis.grid()
    au = H.AU_TO_CM
    plan = evrhip.Plan(basis, type=10, V=ham.potential(q) / au, G0=ham.G0 / au, Ga=ham.Ga_fortran() / au,
                       max_nvec=64, hermitian=True)
    m1, m2, amp = H.control_config(D)
    mplan = evrhip.Plan(basis, type=0, V=H.dipole(q, m1, m2), max_nvec=64)
    ev, vec, _ = plan.davidson(2, block=2, conv_resi=1e-9, want_vectors=True)
    psi0, phi1 = vec[:, 0].astype(complex), vec[:, 1].astype(complex)
    nf = amp.shape[1]
    field = evrhip.control_field(amp, T, nsteps)
    line = {"workload": "optimal control, SG4 D=6 L=5 (%d packed functions), %d fields x %d coefficients, "
                        "T=%g au, %d Chebyshev steps, checkpoints every 32 steps" % (basis.nb_packed, nf,
                                                                                    amp.shape[0], T, nsteps),
            "E0_au": float(ev[0]), "E1_au": float(ev[1])}
    ref = None
    for tag in ("matrix_free", "assembled"):
        if tag == "assembled":
            plan.assemble(evrhip.ASM_FP64)
            mplan.assemble(evrhip.ASM_FP64)
        t0 = time.perf_counter()
        J, g, info = plan.control_gradient(mplan, amp, psi0, phi1, T, nsteps, checkpoint=32, want=("a", "V"))
        wall = time.perf_counter() - t0
        run = {"seconds": round(wall, 3), "seconds_forward": round(info.seconds_forward, 3),
               "seconds_backward": round(info.seconds_backward, 3), "hpsi_columns": info.hpsi,
               "max_order": info.max_order, "J_mean": float(J.mean())}
        if ref is None:
            ref = (J, g)
        else:
            run["max_abs_dJ_vs_matrix_free"] = float(np.abs(J - ref[0]).max())
            run["max_rel_dgrad_a_vs_matrix_free"] = float(np.abs(g["a"] - ref[1]["a"]).max() /
                                                          np.abs(ref[1]["a"]).max())
        psi = np.repeat(psi0[:, None], nf, axis=1)
        for graph in (False, True):
            t0 = time.perf_counter()
            out, _, pinfo = plan.propagate(psi, T / nsteps, nsteps, npoly=200, poly_tol=1e-14, field=field,
                                           mu=mplan, graph=graph)
            run["forward_%s_seconds" % ("graph" if graph else "adaptive")] = round(time.perf_counter() - t0, 3)
        line[tag] = run
        log("%s: %s" % (tag, run))
    print(json.dumps(line), flush=True)


def multi_run(args):
    """64 seeded systems, D uniform in 6..10 and L in 4..6 (uneven basis and
    grid sizes), in atomic units, one wavepacket each (a normalised seeded
    vector), 100 Chebyshev steps of 10 au: all systems as one block-diagonal
    operator (one H|psi> launch sequence per order for all of them) against
    the same propagations one system at a time."""
    import numpy as np
    import evrhip
    from evrhip import hamiltonian as H
    nsys, nsteps, dt = 64, 100, 10.0
    sel = H.SplitMix64(H.config_seed(5), H.FAM_OMEGA).uniform(2 * nsys)
    dims = [6 + min(4, int((u + 1.0) * 2.5)) for u in sel[:nsys]]
    levels = [4 + min(2, int((u + 1.0) * 1.5)) for u in sel[nsys:]]
    plans, lo, hi = [], [], []
    t0 = time.perf_counter()
    for k in range(nsys):
        ham = H.smolyak_config(100 + k, dims[k]).scaled(1.0 / H.AU_TO_CM)
        b = evrhip.Basis.sg4(dims[k], levels[k])
        p = evrhip.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2, hermitian=True)
        p.gen_potential_poly(ham)
        p.assemble(evrhip.ASM_FP64)
        a, z = p.bounds(60)
        pad = 0.02 * (z - a)
        plans.append(p)
        lo.append(a - pad)
        hi.append(z + pad)
    t_setup = time.perf_counter() - t0
    m = evrhip.Multi(plans, max_nvec=2)
    psi = np.concatenate([H.test_vectors(p.nb_packed, 1, 200 + k)[:, 0] for k, p in enumerate(plans)]).astype(complex)
    m.propagate(psi, dt, 2, lo, hi, npoly=200, poly_tol=1e-12)  # warm-up
    t0 = time.perf_counter()
    out, info = m.propagate(psi, dt, nsteps, lo, hi, npoly=200, poly_tol=1e-12)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    diff, orders = 0.0, []
    for k, p in enumerate(plans):
        a, b = m.offsets[k], m.offsets[k + 1]
        ref, _, inf = p.propagate(psi[a:b], dt, nsteps, npoly=200, poly_tol=1e-12, Hmin=lo[k], Hmax=hi[k])
        orders.append(inf.max_order)
        diff = max(diff, float(np.abs(out[a:b] - ref).max()))
    t_seq = time.perf_counter() - t0
    sizes = [int(m.offsets[k + 1] - m.offsets[k]) for k in range(nsys)]
    pool = sum(p.assembled()[1] for p in plans)
    # every batched H|psi> (2 columns: Re, Im) streams all systems' matrices once
    napply = info.hpsi // 2
    line = {"workload": "%d systems, D in 6..10, L in 4..6, one wavepacket each, %d Chebyshev steps of %g au"
                        % (nsys, nsteps, dt),
            "packed_functions_total": m.nb_total, "packed_functions_min": min(sizes),
            "packed_functions_max": max(sizes), "setup_seconds": round(t_setup, 2),
            "batched_seconds": round(t_batch, 3), "one_by_one_seconds": round(t_seq, 3),
            "speedup": round(t_seq / t_batch, 2), "batched_hpsi_columns": info.hpsi,
            "max_order": info.max_order, "orders_one_by_one_max": max(orders),
            "max_abs_diff_vs_one_by_one": diff, "assembled_gbytes": round(pool / 1e9, 2),
            "matrix_stream_gbs": round(pool * napply / t_batch / 1e9, 1),
            "matrix_stream_frac_of_hbm_peak": round(pool * napply / t_batch / 1e9 / PEAK_HBM, 3)}
    print(json.dumps(line), flush=True)


def pyrazine_run(args):
    """The reference's 24D pyrazine propagation test (two coupled states, SG4
    L=2 with a law per mode, 1,502 functions per state; 60 Chebyshev steps of
    0.1 fs, npoly 300, poly_tol 1e-10): GPU wall time adaptive, captured and
    matrix-free, the autocorrelation's distance to the reference's benchmark,
    and the oracle's Chebyshev propagation (C/OpenMP H|psi> per order) on the
    host cores over the first steps, scaled to 60."""
    import numpy as np
    import evrhip
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyrazine import P24 as P
    D, LG, LB, nbA, nbB, nqA, nqB = P.sg4_args()
    b = evrhip.Basis.sg4_modes(D, LG, nbA, nbB, nqA, nqB, LB=LB)
    V11, V22, V12 = P.potentials(b.grid())

    def coupled(assemble):
        p11 = evrhip.Plan(b, type=1, V=V11, F2=P.F2(), max_nvec=2)
        p22 = evrhip.Plan(b, type=1, V=V22, F2=P.F2(), max_nvec=2)
        w = evrhip.Plan(b, type=0, V=V12, max_nvec=2)
        if assemble:
            for p in (p11, p22, w):
                p.assemble(evrhip.ASM_FP64)
        return evrhip.Coupled([[p11, w], [w, p22]])

    n = b.nb_packed
    p0 = np.zeros(2 * n, complex)
    p0[n] = 1.0
    t_ref, ac_ref = P.autocorr_ref()
    out = {}
    for name, assemble, graph in (("assembled", True, False), ("assembled_graph", True, True),
                                  ("matrix_free", False, False)):
        H = coupled(assemble)
        H.propagate(p0, P.DT, 2, npoly=P.NPOLY, poly_tol=P.POLY_TOL, graph=graph)  # warm-up
        t0 = time.perf_counter()
        psi, ac, info = H.propagate(p0, P.DT, P.NSTEPS, npoly=P.NPOLY, poly_tol=P.POLY_TOL,
                                    autocorr=True, graph=graph)
        dt = time.perf_counter() - t0
        out[name] = {"seconds": round(dt, 4), "hpsi_columns": info.hpsi, "max_order": info.max_order,
                     "max_abs_autocorr_diff_vs_reference": float(np.abs(np.r_[1.0, ac[:, 0]] - ac_ref).max())}
        bounds = (info.Hmin, info.Hmax)
        H.close()
    cpu = None
    if not args.no_cpu:
        import oracle
        import propa
        threads = args.cpu_threads or os.cpu_count()
        s = oracle.SG4(D, LG, LB, nbA=nbA, nbB=nbB, nqA=nqA, nqB=nqB)
        ops = [oracle.Op(1, V=V11, F2=P.F2()), oracle.Op(1, V=V22, F2=P.F2()), oracle.Op(0, V=V12)]

        class CoupledOracle:
            def __matmul__(self, w):
                w = np.asarray(w)
                r = np.concatenate([w.real, w.imag], axis=1)
                a = s.oppsi(ops[0], r[:n], nthreads=threads) + s.oppsi(ops[2], r[n:], nthreads=threads)
                c = s.oppsi(ops[2], r[:n], nthreads=threads) + s.oppsi(ops[1], r[n:], nthreads=threads)
                y = np.concatenate([a, c])
                k = w.shape[1]
                return y[:, :k] + 1j * y[:, k:]
        ksteps = P.NSTEPS
        p = p0.reshape(-1, 1).copy()
        t0 = time.perf_counter()
        for _ in range(ksteps):
            p, _, _ = propa.cheby_step(CoupledOracle(), p, P.DT, bounds[0], bounds[1], P.NPOLY, P.POLY_TOL)
        dt = time.perf_counter() - t0
        cpu = {"value": round(dt / ksteps * P.NSTEPS, 3), "unit": "s for 60 steps", "cores": threads,
               "kind": "port",
               "sample": "oracle march_cheby with the oracle's C/OpenMP coupled H|psi> per order, "
                         "first %d of 60 steps (%.2f s wall), scaled" % (ksteps, dt)}
    line = {"workload": "24D pyrazine, 2 coupled states, SG4 L=2 (1502 functions per state), "
                        "60 Chebyshev steps of 0.1 fs, npoly 300, poly_tol 1e-10",
            "gpu": out, "cpu_baseline": cpu}
    if cpu:
        line["gpu_speedup_assembled"] = round(cpu["value"] / out["assembled"]["seconds"], 1)
    print(json.dumps(line), flush=True)


def keo_grid_run(args):
    """Type 10 with G, J and s tabulated per grid point (the reference's grid
    form) on D=6 SG4 L=5: H|psi> of 8 vectors per step on the GPU (FP64, and
    FP32 transforms and grid columns), the same operator through the oracle
    on the host cores.  G = G0 + Ga q of the
    D=6 L=5 Hamiltonian plus a q^2 term, J and s smooth positive functions."""
    import numpy as np
    import torch
    import evrhip
    from evrhip import hamiltonian as H
    D, L, nvec = 6, 5, 8
    ham = H.smolyak_config(1, D)
    b = evrhip.Basis.sg4(D, L)
    q = b.grid()
    Ga = ham.Ga_fortran().reshape(D, D, D, order="F")
    Gq = ham.G0[None] + np.einsum("pk,ijk->pji", q, Ga) + 5.0 * (q ** 2).sum(1)[:, None, None] * np.eye(D)[None]
    J = 1.0 + 0.1 * np.tanh(q @ np.linspace(-0.3, 0.3, D))
    s = 1.0 / np.sqrt(J)
    V = ham.potential(q)
    plan = evrhip.Plan(b, type=10, V=V, G0=ham.G0, max_nvec=nvec)
    plan.set_keo_grid(Gq, J, s)
    X = torch.from_numpy(np.ascontiguousarray(H.test_vectors(b.nb_packed, nvec, 1).T)).cuda()
    Hx = torch.empty_like(X)
    stream = torch.cuda.ExternalStream(plan.stream())
    for _ in range(max(1, args.warmup)):
        plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), nvec)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), nvec)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # the same with FP32 transforms and grid columns
    H32 = torch.empty_like(X)
    for _ in range(max(1, args.warmup)):
        plan.TabOpPsi_device(X.data_ptr(), H32.data_ptr(), nvec, evrhip.FP32)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        plan.TabOpPsi_device(X.data_ptr(), H32.data_ptr(), nvec, evrhip.FP32)
    e1.record(stream)
    torch.cuda.synchronize()
    ms32 = e0.elapsed_time(e1) / args.steps
    # algorithmic HBM bytes of the three grid kernels per vector and grid point:
    # deriv reads psi_G and s, writes D columns; combine reads D columns, D^2
    # G values and J, writes D; out reads D columns, J and s, writes 1
    per_point = 8.0 * ((2 + D) + (D + D * D + 1 + D) + (D + 2 + 1))
    line = {"workload": "type 10 with G, J, s per grid point, SG4 D=6 L=5 (%d functions, %d grid points), "
                        "%d vectors" % (b.nb_packed, b.sum_nq, nvec),
            "ms_per_hpsi_block": round(ms, 3), "hpsi_per_s_per_vector": round(nvec * 1e3 / ms, 1),
            "ms_per_hpsi_block_fp32": round(ms32, 3),
            "fp32_max_rel_diff_vs_fp64": float((H32 - Hx).abs().max() / Hx.abs().max()),
            "grid_kernel_bytes_per_block_gb": round(per_point * b.sum_nq * nvec / 1e9, 3)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = args.cpu_threads or os.cpu_count()
        so = oracle.SG4(D, L)
        op = oracle.Op(10, V=V, Jac=J, s=s, Gq=Gq)
        Xh = X.cpu().numpy().T.copy(order="F")
        ncalls = 0
        t0 = time.perf_counter()
        while True:   # the 8 vectors, repeated for at least 5 s of CPU work
            ref = so.oppsi(op, Xh, nthreads=threads)
            ncalls += 1
            dt = time.perf_counter() - t0
            if dt > 5.0:
                break
        got = Hx.cpu().numpy().T
        line["max_rel_diff_vs_oracle"] = float(np.abs(got - ref).max() / np.abs(ref).max())
        line["cpu_baseline"] = {"value": round(nvec * ncalls / dt, 3), "unit": "Hpsi/s per vector",
                                "cores": threads, "kind": "port",
                                "sample": "oracle grid form (C/OpenMP), %d calls of %d vectors, %.1f s"
                                          % (ncalls, nvec, dt)}
        line["gpu_over_cpu"] = round(line["hpsi_per_s_per_vector"] / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


def transforms_run(args):
    """Basis -> grid and grid -> basis of one vector on the D=9 SG4 L=8 grids
    (sub_PsiBasisRep_TO_GridRep / sub_PsiGridRep_TO_BasisRep): the term
    kernels' B->G and weighted G->B modes, timed by the plan's HIP events
    (the host copies of the 6.4 GB grid vector are outside the kernel times)."""
    import numpy as np
    import evrhip
    from evrhip import hamiltonian as H
    D, L = args.D, args.L
    ham = H.smolyak_config(config_index(D, L), D)
    b = evrhip.Basis.sg4(D, L)
    plan = evrhip.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=1)
    plan.gen_potential_poly(ham)
    x = H.test_vectors(b.nb_packed, 1, 3)[:, 0]
    g = plan.BasisRep_TO_GridRep(x)          # warm-up (allocates the grid buffer)
    plan.GridRep_TO_BasisRep(g)
    plan.set_profiling(True)
    ms = lambda: sum(s[1] for s in plan.kernel_stats() if s[0].startswith("k_term"))
    m0 = ms()
    g = plan.BasisRep_TO_GridRep(x)
    m1 = ms()
    y = plan.GridRep_TO_BasisRep(g)
    m2 = ms()
    # algorithmic work: B->G reads the term's nb_t values and writes nq_t grid
    # values (G->B the reverse); flops of the mode-by-mode contractions
    d = b.ptr.contents
    nbs = np.ctypeslib.as_array(d.nb, shape=(D * b.nlev,))
    nqs = np.ctypeslib.as_array(d.nq, shape=(D * b.nlev,))
    lev = b.term_levels()
    slot = lev + b.nlev * np.arange(D)[None, :]
    tnb, tnq = nbs[slot].astype(float), nqs[slot].astype(float)
    fl = 0.0
    for dd in range(D):
        fl += (2.0 * np.prod(tnq[:, :dd], axis=1) * tnq[:, dd] * tnb[:, dd] * np.prod(tnb[:, dd + 1:], axis=1)).sum()
    nbytes = 8.0 * (b.sum_nb + b.sum_nq)
    line = {"workload": "SG4 D=%d L=%d basis <-> grid, one vector (%d packed functions, %d grid points)"
                        % (D, L, b.nb_packed, b.sum_nq)}
    for tag, t in (("basis_to_grid", m1 - m0), ("grid_to_basis", m2 - m1)):
        line[tag] = {"kernel_ms": round(t, 3), "gbytes": round(nbytes / 1e9, 3),
                     "achieved_gbs": round(nbytes / t / 1e6, 1),
                     "frac_of_hbm_peak": round(nbytes / t / 1e6 / PEAK_HBM, 3),
                     "gflop": round(fl / 1e9, 2), "tflops": round(fl / t / 1e9, 2)}
    # GridTOBasis(BasisTOGrid(x)) = x on the packed space (the Smolyak
    # combination of exact per-term transforms)
    line["roundtrip_max_rel"] = float(np.abs(y - x).max() / np.abs(x).max())
    print(json.dumps(line), flush=True)


def auto_path_run(args):
    """EVR_AUTO_PATH against the two forced paths for blocks of 1 .. --nvec
    vectors on the --D/--L workload (FP64 and FP32): which path the cost model
    takes and what each costs, 5 calls per case timed with HIP events on the
    plan stream."""
    import numpy as np
    import torch
    import evrhip
    from evrhip import hamiltonian as H
    D, L, nmax = args.D, args.L, args.nvec
    cfg = config_index(D, L)
    ham = H.smolyak_config(cfg, D)
    b = evrhip.Basis.sg4(D, L)
    plan = evrhip.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=nmax)
    plan.gen_potential_poly(ham)
    plan.assemble(evrhip.ASM_FP64 | evrhip.ASM_FP32)
    X = torch.from_numpy(np.ascontiguousarray(H.test_vectors(b.nb_packed, nmax, cfg).T)).cuda()
    Hx = torch.empty_like(X)
    s = torch.cuda.ExternalStream(plan.stream())
    rows = []
    for prec, pf in (("fp64", evrhip.FP64), ("fp32", evrhip.FP32)):
        for n in range(1, nmax + 1):
            ms = {}
            for name, f in (("matrix_free", evrhip.MATRIX_FREE), ("assembled", 0), ("auto", evrhip.AUTO_PATH)):
                for _ in range(2):
                    plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), n, f | pf)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(s)
                for _ in range(5):
                    plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), n, f | pf)
                e1.record(s)
                torch.cuda.synchronize()
                ms[name] = round(e0.elapsed_time(e1) / 5, 3)
            log("%s nvec %d: matrix-free %.3f assembled %.3f auto %.3f ms"
                % (prec, n, ms["matrix_free"], ms["assembled"], ms["auto"]))
            rows.append(dict(precision=prec, nvec=n, **ms))
    print(json.dumps({"workload": "SG4 D=%d L=%d type-10 H|psi>, EVR_AUTO_PATH" % (D, L), "ms": rows}), flush=True)


def streamed_run(args):
    """H|psi> on direct-product bases from 3 x 12 to 4 x 20 functions (160,000),
    each one term streamed over global memory in whole-device mode passes;
    type 10 with q-linear G, FP64 and FP32, 5 calls timed with HIP events on
    the plan stream."""
    import numpy as np
    import torch
    import evrhip
    from evrhip import hamiltonian as H
    rows = []
    for D, n, nv in ((3, 12, 32), (3, 16, 32), (3, 21, 32), (3, 30, 32), (3, 40, 32), (4, 20, 32)):
        b = evrhip.Basis.dp([n] * D, [n + 1] * D)
        ham = H.smolyak_config(9, D)
        plan = evrhip.Plan(b, type=10, V=ham.potential(b.grid()), G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=nv)
        X = torch.from_numpy(np.ascontiguousarray(H.test_vectors(b.nb_packed, nv, 3).T)).cuda()
        Hx = torch.empty_like(X)
        s = torch.cuda.ExternalStream(plan.stream())
        ms = {}
        for tag, f in (("fp64", evrhip.FP64), ("fp32", evrhip.FP32)):
            for _ in range(2):
                plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), nv, f)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(5):
                plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), nv, f)
            e1.record(s)
            torch.cuda.synchronize()
            ms[tag] = round(e0.elapsed_time(e1) / 5, 4)
        row = {"basis": "DP %d x %d" % (D, n), "functions": b.nb_packed, "points": b.sum_nq, "nvec": nv,
               "ms_fp64": ms["fp64"], "ms_fp32": ms["fp32"],
               "hpsi_per_s_per_vector_fp64": round(1e3 * nv / ms["fp64"], 1)}
        log("%s (%d functions, %d points), %d vectors: FP64 %.3f ms, FP32 %.3f ms"
            % (row["basis"], b.nb_packed, b.sum_nq, nv, ms["fp64"], ms["fp32"]))
        rows.append(row)
        plan.close()
    print(json.dumps({"workload": "type-10 H|psi> on direct-product bases, LDS-resident and streamed terms",
                      "rows": rows}), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary of this workload (profiles/traffic_sg4_d9l8.md), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_sg4_d9l8.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == kernel:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def asm_bytes(basis, nvec):
    """Algorithmic bytes of one assembled FP64 apply: every term matrix once
    (sum nb_t^2 words, no padding), the packed block in, the term rows out."""
    import numpy as np
    d = basis.ptr.contents
    nbs = np.ctypeslib.as_array(d.nb, shape=(basis.D * basis.nlev,)).astype(np.int64)
    slot = basis.term_levels() + basis.nlev * np.arange(basis.D)[None, :]
    nbt = np.prod(nbs[slot], axis=1).astype(float)
    rows = int((basis.map() >= 0).sum())
    return 8.0 * float((nbt * nbt).sum()) + 8.0 * nvec * (basis.nb_packed + rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nvec", type=int, default=32)
    ap.add_argument("--D", type=int, default=9)
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-davidson", action="store_true",
                    help="skip the FP64 / mixed-precision Davidson levels of the default run")
    ap.add_argument("--no-assembled", action="store_true",
                    help="skip the assembled-operator lines (needs sum nb_t^2 x 12 bytes of HBM)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (scale of the HBM counters)")
    ap.add_argument("--davidson", action="store_true",
                    help="only run the FP64 and FP32-refined Davidson levels and their difference")
    ap.add_argument("--control", action="store_true",
                    help="only run the optimal-control objective, gradient and field-driven propagation")
    ap.add_argument("--multi", action="store_true",
                    help="only run the 64-system batched Chebyshev propagation against one system at a time")
    ap.add_argument("--pyrazine", action="store_true",
                    help="the reference's 24D pyrazine propagation test (two coupled states)")
    ap.add_argument("--keo-grid", action="store_true",
                    help="type 10 with G, J, s per grid point on D=6 SG4 L=5")
    ap.add_argument("--transforms", action="store_true",
                    help="basis <-> grid of one vector on the --D/--L grids, kernel times")
    ap.add_argument("--streamed", action="store_true",
                    help="H|psi> on direct-product bases beyond 8192 functions (streamed terms)")
    ap.add_argument("--auto-path", action="store_true",
                    help="EVR_AUTO_PATH against both forced paths for 1 .. --nvec vectors")
    ap.add_argument("--nev", type=int, default=200)
    ap.add_argument("--conv-resi", type=float, default=1e-3)
    ap.add_argument("--refine-resi", type=float, default=1e-2)
    ap.add_argument("--max-it", type=int, default=400)
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the H|psi> engine runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.davidson:
        davidson_levels(args)
        return
    if args.control:
        control_run(args)
        return
    if args.multi:
        multi_run(args)
        return
    if args.pyrazine:
        pyrazine_run(args)
        return
    if args.keo_grid:
        keo_grid_run(args)
        return
    if args.transforms:
        transforms_run(args)
        return
    if args.auto_path:
        auto_path_run(args)
        return
    if args.streamed:
        streamed_run(args)
        return

    import numpy as np
    import torch
    import evrhip
    from evrhip import hamiltonian as H

    D, L, nvec = args.D, args.L, args.nvec
    cfg_index = config_index(D, L)
    ham = H.smolyak_config(cfg_index, D)
    t0 = time.perf_counter()
    basis = evrhip.Basis.sg4(D, L)
    plan = evrhip.Plan(basis, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=nvec)
    plan.gen_potential_poly(ham)
    log("plan: D=%d L=%d terms=%d nb_packed=%d sum_nq=%d (%.1f s)"
        % (D, L, basis.nterms, basis.nb_packed, basis.sum_nq, time.perf_counter() - t0))

    X = torch.from_numpy(np.ascontiguousarray(H.test_vectors(basis.nb_packed, nvec, cfg_index).T)).cuda()
    Hx = torch.empty_like(X)
    stream = torch.cuda.ExternalStream(plan.stream())

    def step(flags):
        plan.TabOpPsi_device(X.data_ptr(), Hx.data_ptr(), nvec, flags)

    def timed(flags):
        for _ in range(args.warmup):
            step(flags)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        w0 = time.perf_counter()
        for _ in range(args.steps):
            step(flags)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        return e0.elapsed_time(e1) / args.steps, wall

    ms64, wall64 = timed(evrhip.FP64 | evrhip.MATRIX_FREE)
    log("matrix-free fp64: %.3f ms per step (%d vectors)" % (ms64, nvec))
    ms32 = None
    if not args.no_fp32:
        ms32, _ = timed(evrhip.FP32 | evrhip.MATRIX_FREE)
        log("matrix-free fp32: %.3f ms per step" % ms32)

    # per-kernel roofline from HIP events around each term-kernel launch
    plan.set_profiling(True)
    step(evrhip.FP64)
    step(evrhip.FP64)
    torch.cuda.synchronize()
    stats = plan.kernel_stats()
    plan.set_profiling(False)
    dom = max(stats, key=lambda s: s[1])
    name, ms, flops, launches = dom
    achieved = flops / (ms / 1e3) / 1e12
    tot_ms = sum(s[1] for s in stats)
    tot_fl = sum(s[2] for s in stats)
    for s in stats:
        log("  %-16s %8.3f ms/launch  %7.2f TFLOP/s" % (s[0], s[1] / max(s[3], 1),
                                                       s[2] / (s[1] / 1e3) / 1e12 if s[1] else 0))
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64, 4),
                "traffic": measured_traffic(name),
                "kernel": name, "avg_launch_ms": round(ms / launches, 4),
                "all_term_kernels_tflops": round(tot_fl / (tot_ms / 1e3) / 1e12, 3)}

    # assembled per-term operators: build once, then dense MFMA H_t X_t per step
    asm = None
    if not args.no_assembled:
        try:
            t0 = time.perf_counter()
            plan.assemble(evrhip.ASM_FP64 | evrhip.ASM_FP32)
            t_asm = time.perf_counter() - t0
            mask, nbytes = plan.assembled()
            log("assembled: %.1f GB (FP64 + FP32) in %.1f s" % (nbytes / 1e9, t_asm))
            ams64, _ = timed(evrhip.FP64)
            ams32, _ = timed(evrhip.FP32)
            log("assembled fp64: %.3f ms per step, fp32: %.3f ms" % (ams64, ams32))
            plan.set_profiling(True)
            step(evrhip.FP64)
            step(evrhip.FP64)
            torch.cuda.synchronize()
            st = [s for s in plan.kernel_stats() if s[0] == "k_apply_asm"][0]
            plan.set_profiling(False)
            per_ms = st[1] / st[3]
            nbytes_launch = asm_bytes(basis, nvec)
            gbs = nbytes_launch / (per_ms / 1e3) / 1e9
            asm = {"value_fp64": round(nvec / (ams64 / 1e3), 3), "ms_per_step_fp64": round(ams64, 4),
                   "value_fp32": round(nvec / (ams32 / 1e3), 3), "ms_per_step_fp32": round(ams32, 4),
                   "fp32_speedup_vs_fp64": round(ams64 / ams32, 3),
                   "assembly_seconds": round(t_asm, 2), "device_gbytes": round(nbytes / 1e9, 2),
                   "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM,
                                "unit": "GB/s", "frac": round(gbs / PEAK_HBM, 4),
                                "kernel": "k_apply_asm<double,2>", "avg_launch_ms": round(per_ms, 4),
                                "algorithmic_bytes": nbytes_launch,
                                "mfma_tflops": round(st[2] / (st[1] / 1e3) / 1e12, 3)}}
            plan.release_assembled()
        except RuntimeError as e:
            log("assembled operator skipped: %s" % e)
    plan.close()
    del X, Hx
    torch.cuda.empty_cache()

    # the eigenproblem of the metric: 200 lowest levels of the symmetric form,
    # plain FP64 against mixed FP32/FP64 (walls with the operator build added)
    dav = None
    if not args.no_davidson:
        hplan = evrhip.Plan(basis, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=nvec, hermitian=True)
        hplan.gen_potential_poly(ham)
        t_build = 0.0
        if asm is not None:
            t0 = time.perf_counter()
            hplan.assemble(evrhip.ASM_FP64 | evrhip.ASM_FP32)
            t_build = time.perf_counter() - t0
        dav = davidson_pair(hplan, args.nev, nvec, args.conv_resi, args.refine_resi, args.max_it, verbose=False)
        dav["operator_build_seconds"] = round(t_build, 2)
        for tag in ("fp64", "fp32_refined"):
            dav[tag]["seconds_with_build"] = round(dav[tag]["seconds"] + t_build, 3)
        hplan.close()

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or max(1, min(len(os.sched_getaffinity(0)),
                                                 int(os.environ.get("OMP_NUM_THREADS", "1024"))))
        log("cpu baseline: %d threads, full term list" % threads)
        cpu = cpu_baseline(D, L, threads)
        log("cpu: %.4g Hpsi/s per vector" % cpu["value"])

    # headline: the fastest FP64 H|psi> of the plan (assembled operators when
    # they fit in HBM, else the sum-factorised kernels); both paths are reported
    matrix_free = {"value_fp64": round(nvec / (ms64 / 1e3), 3), "ms_per_step_fp64": round(ms64, 4),
                   "value_fp32": None if ms32 is None else round(nvec / (ms32 / 1e3), 3),
                   "ms_per_step_fp32": None if ms32 is None else round(ms32, 4),
                   "roofline": roofline}
    if asm is not None:
        path = "assembled per-term operators (FP64 MFMA, %.0f GB FP64 + FP32 in HBM, built in %.1f s)" % (
            asm["device_gbytes"], asm["assembly_seconds"])
        h64, h32, head_roof = asm["ms_per_step_fp64"], asm["ms_per_step_fp32"], dict(asm["roofline"])
        head_roof["traffic"] = measured_traffic("k_apply_asm")
        head_roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the same workload, "
                                       "profiles/traffic_sg4_d9l8.md (not measured in this run)")
    else:
        path = "sum-factorised term kernels (matrix-free)"
        h64, h32, head_roof = ms64, ms32, roofline
    line = {
        "metric": "Hpsi/s per vector, D=%d SG4 L=%d: FP32-refined vs FP64; max |dE| cm-1" % (D, L),
        "value": round(nvec / (h64 / 1e3), 3),
        "unit": "Hpsi/s per vector",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cubic/quartic potential, q-linear G, random normalised vectors)",
        "config": {"workload": "SG4 D=%d L=%d type-10 H|psi>, %d Smolyak terms, block of %d "
                               "vectors" % (D, L, basis.nterms, nvec),
                   "path": path, "nb_packed": basis.nb_packed, "grid_points": basis.sum_nq,
                   "nvec": nvec},
        "fp32": None if h32 is None else {"value": round(nvec / (h32 / 1e3), 3),
                                          "ms_per_step": round(h32, 4),
                                          "speedup_vs_fp64": round(h64 / h32, 3)},
        "davidson": dav,
        "roofline": head_roof,
        "matrix_free": matrix_free,
        "assembled": asm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
