"""Benchmark of the RieCG rhs on one MI355X.

Workload (the throughput configuration): 160^3 hex-split unit box, each hex
split into 6 tets (24,576,000 tets, 4,173,281 nodes, 28,902,880 edges),
nodes jittered +-10% of the spacing from a fixed seed, RCM-renumbered and
then grouped into compact cells of 2048 nodes (RCM order kept inside); FP64
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


def setup(n, seed=20261202, order="rcm+block"):
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh, renumber_mesh, sedov_state
    t0 = time.time()
    m = box_mesh(n, seed=seed)
    m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    if order == "rcm+block":
        m = renumber_mesh(m, X.mesh_block_order(m["coord"], 2048))
    log("mesh %d^3 + %s: %.1fs" % (n, order, time.time() - t0))
    vol = X.mesh_vol(m["coord"], m["inpoel"])
    edge, dint = X.mesh_domint(m["coord"], m["inpoel"])
    sn, sv, besym = X.mesh_symbc(m["coord"], m["triinpoel"], m["faceset"],
                                 [1, 2, 3, 4, 5, 6])
    U = sedov_state(m["coord"], vol)
    log("domint/vol/symbc: %.1fs (E=%d)" % (time.time() - t0, edge.shape[0]))
    return m, vol, edge, dint, sn, sv, besym, U


def cpu_baseline(n_sample=64, reps=40):
    """The oracle's riemann::rhs (one core, the reference's loop structure and
    superedge grouping) on an n_sample^3 box of the same workload family."""
    import oracle as O
    from xyst_amd.meshgen import box_mesh, sedov_state
    m = box_mesh(n_sample, seed=3)
    vol = O.vol(m["coord"], m["inpoel"])
    edge, dint = O.domint(m["coord"], m["inpoel"])
    se = O.domsuped(m["inpoel"], edge, dint)
    _, _, besym = O.symbc_setup(m["coord"], m["triinpoel"], m["faceset"],
                                [1, 2, 3, 4, 5, 6])
    U = sedov_state(m["coord"], vol)
    G = O.grad(se, m["coord"], m["triinpoel"], U) / vol[:, None]
    cfg = O.cfg()
    t = time.perf_counter()
    for _ in range(reps):
        O.rhs(cfg, se, m["coord"], m["triinpoel"], besym, G, U)
    t = (time.perf_counter() - t) / reps
    E = edge.shape[0]
    return {"value": E / t, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "oracle riemann::rhs on a %d^3 box (%d edges, Sedov "
                      "state), mean of %d evaluations, one thread"
                      % (n_sample, E, reps)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--order", default="rcm+block", choices=["rcm", "rcm+block"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from xyst_amd import riemann as X

    m, vol, edge, dint, sn, sv, besym, U = setup(args.n, order=args.order)
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

    # one full RK3 step (dt + 3 x (grad, rhs, update, BCs)) for context
    dUn = torch.empty_like(dU)
    dt = ctx.dt_dev(dU, 0.5)
    ev0.record(stream)
    for s in range(3):
        ctx.stage_dev(s, dt, dU, dUn, dG, dR)
    ev1.record(stream)
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1)

    out = {
        "metric": "rhs edges/s (% HBM peak)",
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
                   "boundary_faces": Fb, "flux": "rusanov", "ordering": args.order,
                   "state": "sedov, rho=1, p=1e-6, E=1 within r<0.02"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None},
        "riecg_step_ms": step_ms,
        "wall_s_timed": wall,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
