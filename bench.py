"""Benchmark of the RieCG rhs on one MI355X.

Workload (the throughput configuration): 160^3 hex-split unit box, each hex
split into 6 tets (24,576,000 tets, 4,173,281 nodes, 28,902,880 edges),
nodes jittered +-10% of the spacing from a fixed seed, RCM-renumbered and
then grouped into compact cells of 1024 nodes (RCM order kept inside); FP64
Sedov blast state (energy 1.0 deposited in the nodes within r < 0.02 of the
centre, rho = 1, p = 1e-6 elsewhere), gamma 1.4, Rusanov flux, symmetry walls.

A step is one riemann::rhs evaluation (domain edges + boundary faces) on the
HBM-resident state and gradients.  value = edges / second.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 160]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def rhs_bytes(E, N, Fb):
    """Algorithmic bytes of one riemann::rhs: per edge two int32 ids + 3 FP64
    D; per node coordinates (24), U (40), G (120) read and R (40) written;
    per boundary face three int32 ids and three symmetry bytes."""
    return 32 * E + 224 * N + 15 * Fb


def setup(n, seed=20261202, order="rcm+block", cell=1024):
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh, renumber_mesh, sedov_state
    t0 = time.time()
    m = box_mesh(n, seed=seed)
    m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    if order == "rcm+block":
        m = renumber_mesh(m, X.mesh_block_order(m["coord"], cell))
    log("mesh %d^3 + %s: %.1fs" % (n, order, time.time() - t0))
    vol = X.mesh_vol(m["coord"], m["inpoel"])
    edge, dint = X.mesh_domint(m["coord"], m["inpoel"])
    sn, sv, besym = X.mesh_symbc(m["coord"], m["triinpoel"], m["faceset"],
                                 [1, 2, 3, 4, 5, 6])
    U = sedov_state(m["coord"], vol)
    log("domint/vol/symbc: %.1fs (E=%d)" % (time.time() - t0, edge.shape[0]))
    return m, vol, edge, dint, sn, sv, besym, U


_POISSON = {}


def poisson_setup(n, seed=20261202, modes=8):
    """poisson_system(n) once per size (the Poisson and learned sections use
    the same systems)."""
    if (n, seed, modes) not in _POISSON:
        _POISSON[(n, seed, modes)] = poisson_system(n, seed, modes)
    return _POISSON[(n, seed, modes)]


def poisson_system(n, seed=20261202, modes=8):
    """ChoCG pressure-Poisson system on an n^3 box: A = -K (ChoCG::prelhs),
    b = the edge-based divergence sum_e D_e . (u_p + u_q) (chorin::div's
    domain term) of a smooth velocity field of `modes` seeded Fourier modes of
    unit amplitude, scaled to ||b|| = 1e4; Dirichlet p = 0 on x = 0, natural
    (Neumann) elsewhere."""
    from xyst_amd import cg as CG
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh, renumber_mesh
    t0 = time.time()
    m = box_mesh(n, seed=seed)
    m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    coord = m["coord"]
    N = coord.shape[1]
    A = CG.CSR.poisson(coord, m["inpoel"])
    rng = np.random.default_rng(seed)
    vel = np.zeros((3, N))
    for _ in range(modes):
        k = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        amp = rng.standard_normal(3)
        amp /= np.linalg.norm(amp)
        arg = 2 * np.pi * (k[0] * coord[0] + k[1] * coord[1] + k[2] * coord[2])
        for j in range(3):
            vel[j] += amp[j] * np.sin(arg + ph[j])
    edge, dint = X.mesh_domint(coord, m["inpoel"])
    p, q = edge[:, 0].astype(np.int64), edge[:, 1].astype(np.int64)
    f = ((vel[:, p] + vel[:, q]) * dint.T).sum(0)
    b = np.bincount(q, f, N) - np.bincount(p, f, N)
    # scaled to ||b|| = 1e4: the reference's CG stops once |(p,q)| falls below
    # DBL_EPSILON in absolute terms, which an O(h^2) right-hand side reaches
    # before a relative residual of 1e-10
    b *= 1e4 / np.linalg.norm(b)
    dirnodes = np.where(coord[0] < 1e-12)[0]
    log("poisson %d^3: N=%d nnz=%d, %d Dirichlet nodes, %.1fs"
        % (n, N, A.ja.size, dirnodes.size, time.time() - t0))
    return A, b, {int(i): [(1, 0.0)] for i in dirnodes}


def bc_residual(A, b, dirnodes):
    """||b' - A' x|| / ||b'|| of the system the solvers solve: the Dirichlet
    rows and columns of A replaced by the identity and b' = 0 there (the
    Dirichlet values are 0), recomputed in FP64 on the host with scipy from
    the solution the device returns -- independent of every solver's own
    residual."""
    import scipy.sparse as sp
    n = A.nrow
    M = sp.csr_matrix((A.a, A.ja, A.ia), shape=(n, n))
    dn = np.asarray(dirnodes, np.int64)
    bb = b.copy()
    bb[dn] = 0.0
    nb = np.linalg.norm(bb)

    def residual(x):
        xm = x.copy()
        xm[dn] = 0.0
        r = bb - M @ xm
        r[dn] = -x[dn]
        return float(np.linalg.norm(r) / nb)
    return residual


def poisson_bench(n, stream, reps=2):
    """ms per pressure-Poisson solve to 1e-10: FP64 Jacobi-PCG (the
    reference algorithm on the device), the FP32-inner solve with FP64
    residual refinement (restarted, and as one FP32 recurrence with FP64
    reliable updates), FP64 flexible CG with an FP32 inner CG, and FP64
    flexible CG with one FP32 smoothed-aggregation multigrid V-cycle per
    iteration.  Timed: setup (r = b - Ax, diagonal, norms) and the
    iterations, with A and b resident on the device; the multigrid
    hierarchy is built on the host once per matrix and Dirichlet set (as
    ChoCG's matrix is), its build time is reported beside, not in, the
    per-solve time.  Every solver reports the true relative residual
    recomputed on the host (bc_residual) and the relative error of its
    solution against the FP64 Jacobi-PCG solution."""
    import torch
    from xyst_amd import cg as CG
    from xyst_amd import riemann as X
    A, b, dirbc = poisson_setup(n)
    residual = bc_residual(A, b, sorted(dirbc))
    ctx = X.Context(0, stream=stream)
    cg = CG.ConjugateGradients(A, ctx=ctx)
    x0 = np.zeros(A.nrow)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    out = {"n": n, "nodes": int(A.nrow), "nnz": int(A.ja.size), "tol": 1e-10}
    for name in ("fp64_jacobi_pcg", "mixed_fp32_fp64", "mixed_fp32_reliable",
                 "fcg_fp64_cg32_inner", "fcg_fp64_amg"):
        ms = []
        for _ in range(reps):
            cg.init(x0, b, dirbc=dirbc, apply=True, pc="jacobi")
            torch.cuda.synchronize()
            ev0.record(stream)
            if name == "fp64_jacobi_pcg":
                cg.setup()
                cg.solve(100000, 1e-10)
            elif name == "mixed_fp32_fp64":
                cg.solve_mixed(100000, 1e-10, inner_tol=1e-5, inner_maxit=5000)
            elif name == "mixed_fp32_reliable":
                cg.solve_reliable(100000, 1e-10, delta=0.1)
            elif name == "fcg_fp64_cg32_inner":
                cg.solve_flexible(100000, 1e-10, pc="cg32", inner=30)
            else:
                cg.solve_flexible(100000, 1e-10, pc="amg")
            ev1.record(stream)
            torch.cuda.synchronize()
            ms.append(ev0.elapsed_time(ev1))
        x = cg.solution()
        if name == "fp64_jacobi_pcg":
            x_ref = x
        out[name] = {"ms": min(ms), "iterations": int(cg.it),
                     "converged": bool(cg.converged),
                     "solver_normr_over_normb": cg.normr / cg.normb,
                     "true_residual_rel": residual(x),
                     "err_vs_fp64_pcg_rel": float(np.linalg.norm(x - x_ref) /
                                                  np.linalg.norm(x_ref))}
        if name == "mixed_fp32_fp64":
            out[name]["outer"] = int(cg.nouter)
        if name == "mixed_fp32_reliable":
            out[name]["reliable_updates"] = int(cg.nupdates)
        if name == "fcg_fp64_cg32_inner":
            out[name]["inner_iterations"] = int(cg.inner_total)
        if name == "fcg_fp64_amg":
            lv, oc, sms = cg.amg_info()
            out[name].update({"first_solve_ms_with_setup": ms[0], "levels": lv,
                              "operator_complexity": oc, "host_setup_ms": sms})
        log("%s: %.2f ms, %d it, %s, true residual %.2e, error vs FP64 PCG %.2e"
            % (name, min(ms), cg.it, cg.converged, out[name]["true_residual_rel"],
               out[name]["err_vs_fp64_pcg_rel"]))
    out["speedup_mixed"] = out["fp64_jacobi_pcg"]["ms"] / out["mixed_fp32_fp64"]["ms"]
    out["speedup_reliable"] = out["fp64_jacobi_pcg"]["ms"] / out["mixed_fp32_reliable"]["ms"]
    out["speedup_fcg_cg32"] = out["fp64_jacobi_pcg"]["ms"] / out["fcg_fp64_cg32_inner"]["ms"]
    out["speedup_fcg_amg"] = out["fp64_jacobi_pcg"]["ms"] / out["fcg_fp64_amg"]["ms"]
    ctx.close()
    return out


def learned_bench(stream, eval_sizes, train_n=64, nsys=36, pairs=600, seed=7):
    """The learned graph preconditioner: trained on the device on residual/
    error pairs of Jacobi-PCG solves of ChoCG pressure systems on a
    train_n^3 box (seeded random smooth right-hand sides, up to `pairs`
    pairs per system), then applied inside FP64 flexible CG on the
    pressure-Poisson systems of the Poisson section (eval_sizes), against
    FP64 Jacobi-PCG on the same system: iterations, ms and the true
    residual of each.  FCG falls back to Jacobi for the rest of a solve once
    the residual has not improved for 20 iterations."""
    import torch
    from xyst_amd import cg as CG
    from xyst_amd import gnn
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh, renumber_mesh
    t0 = time.time()
    m = box_mesh(train_n, seed=seed)
    m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    coord = m["coord"]
    N = coord.shape[1]
    A = CG.CSR.poisson(coord, m["inpoel"])
    dirbc = {int(i): [(1, 0.0)] for i in np.where(coord[0] < 1e-12)[0]}
    ctx = X.Context(0, stream=stream)
    cg = CG.ConjugateGradients(A, ctx=ctx)
    cg.init(np.zeros(N), np.ones(N), dirbc=dirbc, apply=True, pc="jacobi")
    net = gnn.LearnedPreconditioner(cg, seed=seed)
    rng = np.random.default_rng(seed)
    x, y, z = coord
    total, losses = 0, []
    for k in range(nsys):
        b = np.zeros(N)
        for _ in range(4):
            kx, ky, kz = rng.integers(1, 4, size=3)
            b += rng.standard_normal() * np.sin(np.pi * (kx * x + rng.uniform())) * \
                np.cos(np.pi * ky * y) * np.cos(np.pi * (kz * z + rng.uniform()))
        b *= 1e4 / np.linalg.norm(b)
        cg.init(np.zeros(N), b, dirbc=dirbc, apply=True, pc="jacobi")
        L, npairs = net.train_system(stride=1, max_pairs=pairs, batch=8, lr=1e-3)
        total += npairs
        losses.append(L)
    params = net.params()
    train_s = time.time() - t0
    log("learned: trained on %d pairs of %d systems (%d^3) in %.1fs, loss %.4f -> %.4f"
        % (total, nsys, train_n, train_s, losses[0], losses[-1]))
    ctx.close()
    out = {"train_n": train_n, "train_pairs": int(total), "train_systems": nsys,
           "train_s": train_s, "loss_first_system": losses[0],
           "loss_last_system": losses[-1]}
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    for n in eval_sizes:
        A, b, dirbc = poisson_setup(n)
        residual = bc_residual(A, b, sorted(dirbc))
        ctx = X.Context(0, stream=stream)
        cg = CG.ConjugateGradients(A, ctx=ctx)
        net = gnn.LearnedPreconditioner(cg, params=params)
        row = {}
        for name in ("fp64_jacobi_pcg", "fcg_fp64_learned"):
            cg.init(np.zeros(A.nrow), b, dirbc=dirbc, apply=True, pc="jacobi")
            torch.cuda.synchronize()
            ev0.record(stream)
            if name == "fp64_jacobi_pcg":
                cg.setup()
                cg.solve(100000, 1e-10)
                maxit = 3 * cg.it
            else:
                cg.solve_flexible(maxit, 1e-10, pc="gnn")
            ev1.record(stream)
            torch.cuda.synchronize()
            row[name] = {"ms": ev0.elapsed_time(ev1), "iterations": int(cg.it),
                         "converged": bool(cg.converged),
                         "true_residual_rel": residual(cg.solution())}
            if name == "fcg_fp64_learned":
                row[name]["jacobi_fallback_applications"] = int(cg.inner_total)
        row["speedup_learned"] = row["fp64_jacobi_pcg"]["ms"] / row["fcg_fp64_learned"]["ms"]
        log("learned %d^3: PCG %d it %.1f ms; FCG+network %d it %.1f ms"
            % (n, row["fp64_jacobi_pcg"]["iterations"], row["fp64_jacobi_pcg"]["ms"],
               row["fcg_fp64_learned"]["iterations"], row["fcg_fp64_learned"]["ms"]))
        out["eval_%d" % n] = row
        ctx.close()
        torch.cuda.empty_cache()
    return out


def kozak_rhs_bytes(T, N):
    """Algorithmic bytes of one kozak::rhs: per tet four int32 ids; per node
    coordinates (24), U (40) read and R (40) written."""
    return 16 * T + 104 * N


def kozcg_step_bytes(T, N):
    """Algorithmic bytes of one KozCG FCT step done as the three element
    sweeps its data dependences require, each reading the tets (16 B) and the
    node data it needs once: (1) rhs + antidiffusive contributions: U, x in,
    R and P+/- (120) out; (2) allowed bounds: R, P+/-, U, vol in, Q+/- (80)
    and the low-order solution (40) out; (3) limiter + solve: U, x, Q+/-,
    low-order solution, P+/-, vol in, U and Un out."""
    return 48 * T + (64 + 120) * N + (168 + 120) * N + (312 + 80) * N


def kozcg_bench(ctx, m, vol, sn, sv, stream, steps):
    """kozak::rhs and the KozCG FCT step (reference defaults: fctdif 1,
    no clipping, no system coupling) on the same Sedov state and mesh."""
    import torch
    from xyst_amd import kozak as K
    from xyst_amd.meshgen import sedov_state
    t0 = time.time()
    K.upload(ctx, m["coord"], m["inpoel"])
    K.setup(ctx, vol, sn, sv, fct=True, fctdif=1.0, fctclip=False, fctsys=())
    log("kozcg layout: %.1fs" % (time.time() - t0))
    U = torch.from_numpy(sedov_state(m["coord"], vol)).cuda()
    R = torch.empty_like(U)
    Un = torch.empty_like(U)
    N, T = U.shape[0], m["inpoel"].shape[0]
    dt = K.dt_dev(ctx, U, 0.5)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        K.rhs_dev(ctx, dt, U, R)
    ev0.record(stream)
    for _ in range(steps):
        K.rhs_dev(ctx, dt, U, R)
    ev1.record(stream)
    torch.cuda.synchronize()
    rhs_ms = ev0.elapsed_time(ev1) / steps
    U0 = U.clone()
    for _ in range(3):
        K.step_dev(ctx, dt, U, Un)
    U.copy_(U0)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        K.step_dev(ctx, dt, U, Un)
    ev1.record(stream)
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / steps
    rb, sb = kozak_rhs_bytes(T, N), kozcg_step_bytes(T, N)
    out = {"tets": T, "nodes": N, "rhs_ms": rhs_ms, "rhs_tets_per_s": T / (rhs_ms * 1e-3),
           "rhs_algorithmic_GBs": rb / (rhs_ms * 1e-3) / 1e9,
           "rhs_frac_hbm": rb / (rhs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "step_ms": step_ms, "step_algorithmic_GBs": sb / (step_ms * 1e-3) / 1e9,
           "step_frac_hbm": sb / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "fct": "fctdif 1, fctclip off, fctsys none", "dt": dt}
    log("kozcg: rhs %.3f ms, step %.3f ms" % (rhs_ms, step_ms))
    del U, R, Un, U0
    return out


def chocg_bench(n, stream, steps, p_pc="jacobi"):
    """The ChoCG time step (explicit momentum + pressure projection) on the
    Taylor-Green vortex in [0, pi]^3 with free-slip (symmetry) walls, which
    its symmetries make equivalent to the periodic 2 pi box: u = sin x cos y
    cos z, v = -cos x sin y cos z, w = 0; Re 1600 (mu = 1/1600), damp4 flux
    with stab, rk 1, cfl 0.5, pressure Jacobi-PCG to 1e-6 (warm-started),
    hydrostat node 0.  Times `steps` steps after the start-up projections."""
    import torch
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    from xyst_amd.cg import CSR, ConjugateGradients
    from xyst_amd.meshgen import box_mesh, renumber_mesh
    t0 = time.time()
    m = box_mesh(n, seed=20261202, length=(np.pi, np.pi, np.pi))
    m = renumber_mesh(m, R.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    m = renumber_mesh(m, R.mesh_block_order(m["coord"], 2048))
    coord, inpoel = m["coord"], m["inpoel"]
    N = coord.shape[1]
    vol = R.mesh_vol(coord, inpoel)
    edge, dint = X.mesh_domint(coord, inpoel)
    sn, sv, _ = R.mesh_symbc(coord, m["triinpoel"], m["faceset"], [1, 2, 3, 4, 5, 6])
    ctx = R.Context(0, stream=stream)
    sup, ints = R.edges_as_superedges(edge, dint)
    X.upload(ctx, coord, sup, ints, m["triinpoel"])
    A = CSR.poisson(coord, inpoel)
    cg = ConjugateGradients(A, ctx=ctx)
    X.step_setup(ctx, vol, X.config("damp4", stab=True, mu=1.0 / 1600.0), cfl=0.5, rk=1,
                 p_pc=p_pc, p_maxit=5000, p_tol=1e-6, symnodes=sn, symnorms=sv,
                 pdirnodes=np.array([0], np.int64), pdirvals=np.zeros(1))
    x, y, z = coord
    U0 = np.stack([np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z),
                   np.zeros(N)], 1)
    U = torch.tensor(U0, device="cuda")
    P = torch.zeros(N, dtype=torch.float64, device="cuda")
    Un, dP = torch.empty_like(U), torch.empty_like(P)
    pit0 = X.initial_dev(ctx, U, P)
    torch.cuda.synchronize()
    log("chocg tgv %d^3 setup + start-up: %.1fs (pressure its %s)"
        % (n, time.time() - t0, list(pit0)))
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    its, ms = [], []
    for _ in range(steps):
        dt = X.dt_dev(ctx, U)
        ev0.record(stream)
        its.append(X.step_dev(ctx, dt, U, Un, P, dP))
        ev1.record(stream)
        torch.cuda.synchronize()
        ms.append(ev0.elapsed_time(ev1))
    nnz = int(A.ja.size)
    out = {"n": n, "nodes": N, "edges": int(edge.shape[0]), "nnz": nnz,
           "ms_per_step": float(np.mean(ms)), "ms_steps": ms, "pressure_its": its,
           "startup_pressure_its": list(pit0), "dt": dt,
           "config": "TGV in [0,pi]^3, free-slip walls, Re 1600, damp4+stab, rk1, "
                     "cfl 0.5, pressure solve to 1e-6 with %s"
                     % ("Jacobi-PCG" if p_pc == "jacobi" else "FCG + AMG")}
    log("chocg (%s): %.2f ms/step, pressure its %s" % (p_pc, out["ms_per_step"], its))
    ctx.close()
    del U, P, Un, dP
    return out


def copy_bandwidth(stream, nbytes=1 << 30, reps=10):
    """Measured device-to-device copy rate (read + write bytes per second) of
    an nbytes buffer (torch copy_, i.e. the HIP runtime's device copy),
    timed with HIP events: the attainable HBM bandwidth the kernels are also
    compared with."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        b.copy_(a)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    del a, b
    return {"GBs": 2.0 * nbytes / (ms * 1e-3) / 1e9, "bytes_copied": nbytes,
            "ms": ms, "frac_of_peak": 2.0 * nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


RHS_SOURCES = ("xyst_amd/csrc/xg_riemann.hip", "xyst_amd/csrc/xg_flux.hpp",
               "xyst_amd/csrc/xg_blocks.hpp", "xyst_amd/csrc/xg_layout.cpp",
               "xyst_amd/csrc/xg_layout.hpp")


def rhs_sources_sha16():
    """Hash of the sources the rhs kernels and their layout are built from."""
    import hashlib
    h = hashlib.sha256()
    for f in RHS_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def read_traffic(path):
    """HBM bytes per rhs launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by what the counter reads for a copy
    of known size), as recorded in the profile summary together with the hash
    of the rhs sources it was measured on; None if absent or if the sources
    have changed since (the measurement would be stale)."""
    if not os.path.exists(path):
        return None
    traffic, sha = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_rhs:"):
            traffic = float(line.split(":", 1)[1])
        if line.startswith("rhs_sources_sha16:"):
            sha = line.split(":", 1)[1].strip()
    if sha != rhs_sources_sha16():
        log("rhs traffic in %s was measured on other rhs sources: not reported"
            % os.path.basename(path))
        return None
    return traffic


def cpu_threads():
    """Host threads for the all-cores baseline: this process's CPU share
    (affinity), capped by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(m, edge, dint, besym, U, vol, reps=2, reps_mt=10):
    """The oracle's riemann::rhs (the reference's loop structure and
    superedge grouping) on the bench's own mesh and state (the same
    renumbered 160^3 box, Sedov state, symmetry walls): one thread, and all
    of this process's cores with one contiguous superedge partition per
    thread and the per-node sum of the partitions (the reference's one chare
    per core).  `value` is the all-cores rate.  The superedges are grouped by
    the oracle from the bench's edge list."""
    import oracle as O
    t0 = time.time()
    se = O.domsuped(m["inpoel"], edge, dint)
    G = O.grad(se, m["coord"], m["triinpoel"], U) / vol[:, None]
    cfg = O.cfg()
    log("cpu baseline setup: %.1fs" % (time.time() - t0))
    t = time.perf_counter()
    for _ in range(reps):
        O.rhs(cfg, se, m["coord"], m["triinpoel"], besym, G, U)
    t1 = (time.perf_counter() - t) / reps
    T = cpu_threads()
    O.rhs_mt(cfg, se, m["coord"], m["triinpoel"], besym, G, U, T)
    t = time.perf_counter()
    for _ in range(reps_mt):
        O.rhs_mt(cfg, se, m["coord"], m["triinpoel"], besym, G, U, T)
    tT = (time.perf_counter() - t) / reps_mt
    E = edge.shape[0]
    return {"value": E / tT, "unit": "edges/s", "cores": T, "kind": "port",
            "sample": "oracle riemann::rhs on the bench's %d-edge mesh and Sedov state: "
                      "%d threads, one superedge partition each, mean of %d evaluations; "
                      "one thread, mean of %d evaluations" % (E, T, reps_mt, reps),
            "single_core_value": E / t1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--order", default="rcm+block", choices=["rcm", "rcm+block"])
    ap.add_argument("--block-cell", type=int, default=1024,
                    help="nodes per cell of the RCB block order (rcm+block)")
    ap.add_argument("--poisson-n", default="128,192",
                    help="box sizes of the pressure-Poisson solves, comma separated "
                         "(0: skip)")
    ap.add_argument("--kozcg-steps", type=int, default=20,
                    help="timed kozak::rhs / KozCG steps on the same mesh (0: skip)")
    ap.add_argument("--chocg-n", type=int, default=128,
                    help="box size of the ChoCG Taylor-Green steps (0: skip)")
    ap.add_argument("--chocg-steps", type=int, default=5,
                    help="timed ChoCG Taylor-Green steps per pressure solver")
    ap.add_argument("--chocg-pc", default="both", choices=["both", "jacobi", "amg"],
                    help="pressure solver(s) of the ChoCG steps")
    ap.add_argument("--learned-n", type=int, default=64,
                    help="box size the learned preconditioner is trained on; it is "
                         "evaluated on the Poisson sizes (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, copy a 512 MiB buffer 3 times "
                         "(the known-size kernel that calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from xyst_amd import riemann as X

    m, vol, edge, dint, sn, sv, besym, U = setup(args.n, order=args.order, cell=args.block_cell)
    N, E, Fb = m["coord"].shape[1], edge.shape[0], m["triinpoel"].shape[0]
    stream = torch.cuda.current_stream()
    ctx = X.Context(0, flux="rusanov", stream=stream)
    sup, ints = X.edges_as_superedges(edge, dint)
    t0 = time.time()
    ctx.upload_mesh(m["coord"], sup, ints, m["triinpoel"], besym)
    ctx.riecg_setup(vol, sn, sv)
    log("device layout: %.1fs, %.2f GB" % (time.time() - t0, ctx.device_bytes() / 1e9))

    dU = torch.from_numpy(U).cuda()
    dG = torch.empty((N, 15), dtype=torch.float64, device="cuda")
    dR = torch.empty_like(dU)
    dvol = torch.from_numpy(vol).cuda()
    ctx.grad_dev(dU, dG)
    dG /= dvol[:, None]
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        ctx.rhs_dev(dG, dU, dR)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ctx.rhs_dev(dG, dU, dR)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    ms = ev0.elapsed_time(ev1) / args.steps
    value = E / (ms * 1e-3)
    B = rhs_bytes(E, N, Fb)
    achieved = B / (ms * 1e-3) / 1e9

    # full RK3 steps (dt + 3 x (grad, rhs, update, BCs)) for context, the
    # mean of a few
    dUn = torch.empty_like(dU)
    rk_reps = 5
    ev0.record(stream)
    for _ in range(rk_reps):
        dt = ctx.dt_dev(dU, 0.5)
        for s in range(3):
            ctx.stage_dev(s, dt, dU, dUn, dG, dR)
    ev1.record(stream)
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / rk_reps

    out = {
        "metric": "rhs edges/s (% HBM peak); ms per Poisson solve to 1e-10 vs FP64 Jacobi-PCG",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "riecg_rhs_sedov_box%d" % args.n, "n": args.n,
                   "nodes": N, "edges": E, "tets": int(m["inpoel"].shape[0]),
                   "boundary_faces": Fb, "flux": "rusanov",
                   "ordering": args.order + ("" if args.order == "rcm" else
                                             " (cells of %d nodes)" % args.block_cell),
                   "state": "sedov, rho=1, p=1e-6, E=1 within r<0.02"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None},
        "riecg_step_ms": step_ms,
        "wall_s_timed": wall,
    }
    out["hbm_copy"] = copy_bandwidth(stream)
    out["hbm_copy"]["rhs_frac_of_copy"] = achieved / out["hbm_copy"]["GBs"]
    traffic = read_traffic(os.path.join(ROOT, "profiles", "rhs_box160_traffic.txt"))
    if traffic is not None and args.n == 160:
        out["roofline"]["traffic"] = traffic
    if args.copy_probe:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.kozcg_steps > 0:
        out["kozcg"] = kozcg_bench(ctx, m, vol, sn, sv, stream, args.kozcg_steps)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(m, edge, dint, besym, U, vol)
    psizes = [int(v) for v in str(args.poisson_n).split(",") if int(v) > 0]
    if psizes:
        ctx.close()
        del dG, dR, dU, dUn
        torch.cuda.empty_cache()
    for k, pn in enumerate(psizes):
        torch.cuda.empty_cache()
        out["poisson" if k == 0 else "poisson_%d" % pn] = poisson_bench(pn, stream)
    if args.learned_n > 0 and psizes:
        torch.cuda.empty_cache()
        out["learned"] = learned_bench(stream, psizes, train_n=args.learned_n)
    if args.chocg_n > 0 and args.chocg_pc in ("both", "jacobi"):
        torch.cuda.empty_cache()
        out["chocg"] = chocg_bench(args.chocg_n, stream, args.chocg_steps)
    if args.chocg_n > 0 and args.chocg_pc in ("both", "amg"):
        torch.cuda.empty_cache()
        out["chocg_amg"] = chocg_bench(args.chocg_n, stream, args.chocg_steps, p_pc="amg")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
