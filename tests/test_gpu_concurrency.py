"""Reentrancy (needs an MI355X): one context and one stream per partition,
as the reference runs one chare per partition.  Two host threads, each with
its own context, stream and mesh, run the RieCG stage and the KozCG step at
the same time; every result equals, bit for bit, the one the same context
produces alone."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _partition(n, seed):
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh, sod_state
    import torch
    m = box_mesh(n, seed=seed)
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    vol = X.mesh_vol(coord, inpoel)
    edge, dint = X.mesh_domint(coord, inpoel)
    sup, ints = X.edges_as_superedges(edge, dint)
    sn, sv, besym = X.mesh_symbc(coord, tri, m["faceset"], [1, 2, 3, 4, 5, 6])
    stream = torch.cuda.Stream()
    ctx = X.Context(0, stream=stream)
    ctx.upload_mesh(coord, sup, ints, tri, besym)
    ctx.riecg_setup(vol, sn, sv)
    K.upload(ctx, coord, inpoel)
    K.setup(ctx, vol, sn, sv)
    U0 = sod_state(coord)
    return ctx, stream, U0


def _run(ctx, stream, U0, steps, out):
    import torch
    from xyst_amd import kozak as K
    with torch.cuda.stream(stream):
        U = torch.tensor(U0, device="cuda")
        Un = torch.empty_like(U)
        G = torch.empty((U.shape[0], 15), dtype=torch.float64, device="cuda")
        R = torch.empty_like(U)
        for _ in range(steps):
            dt = ctx.dt_dev(U, 0.5)
            for s in range(3):
                ctx.stage_dev(s, dt, U, Un, G, R)
        V = torch.tensor(U0, device="cuda")
        Vn = torch.empty_like(V)
        for _ in range(steps):
            K.step_dev(ctx, K.dt_dev(ctx, V, 0.5), V, Vn)
        stream.synchronize()
        out.append((U.cpu().numpy(), V.cpu().numpy()))


def test_two_partitions_in_two_threads():
    _say("concurrency: two contexts, two streams, two threads")
    parts = [_partition(20, 3), _partition(24, 4)]
    alone = []
    for ctx, stream, U0 in parts:
        res = []
        _run(ctx, stream, U0, 5, res)
        alone.append(res[0])
    for rep in range(3):
        outs = [[], []]
        th = [threading.Thread(target=_run, args=(ctx, stream, U0, 5, outs[i]))
              for i, (ctx, stream, U0) in enumerate(parts)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i in range(2):
            assert len(outs[i]) == 1, "thread %d failed" % i
            assert np.array_equal(outs[i][0][0], alone[i][0]), (rep, i, "RieCG")
            assert np.array_equal(outs[i][0][1], alone[i][1]), (rep, i, "KozCG")
    for ctx, _, _ in parts:
        ctx.close()
