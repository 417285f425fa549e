"""Device memory over context lifetimes (needs an MI355X): repeated mesh
uploads and setups on one context replace their previous allocations, and
destroying a context returns everything it allocated, for the RieCG,
KozCG, ChoCG (with the multigrid pressure solver) and CG paths."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _free():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return torch.cuda.mem_get_info()[0]


def _one_lifetime(m, vol, sn, sv, besym, sup, ints, chsup, chints, A, repeat):
    import torch
    from xyst_amd import chorin as CH
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    from xyst_amd.cg import ConjugateGradients
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    n = coord.shape[1]
    ctx = X.Context(0)
    for _ in range(repeat):     # setups replace what they allocated before
        ctx.upload_mesh(coord, sup, ints, tri, besym)
        ctx.riecg_setup(vol, sn, sv)
        K.upload(ctx, coord, inpoel)
        K.setup(ctx, vol, sn, sv)
    U = torch.zeros((n, 5), dtype=torch.float64, device="cuda")
    U[:, 0] = 1.0
    U[:, 4] = 2.5
    Un = torch.empty_like(U)
    G = torch.empty((n, 15), dtype=torch.float64, device="cuda")
    R = torch.empty_like(U)
    dt = ctx.dt_dev(U, 0.5)
    ctx.stage_dev(0, dt, U, Un, G, R)
    K.step_dev(ctx, K.dt_dev(ctx, U, 0.5), U, Un)
    for _ in range(repeat):
        CH.upload(ctx, coord, chsup, chints, tri)
        cg = ConjugateGradients(A, ctx=ctx)
        CH.step_setup(ctx, vol, CH.config("damp2", stab=True, mu=0.01), cfl=0.5, rk=1,
                      p_pc="amg", p_maxit=200, p_tol=1e-6, symnodes=sn, symnorms=sv,
                      pdirnodes=np.array([0], np.int64), pdirvals=np.zeros(1))
    V = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    V[:, 0] = torch.tensor(np.sin(coord[1]), device="cuda")
    P = torch.zeros(n, dtype=torch.float64, device="cuda")
    CH.initial_dev(ctx, V, P)
    CH.step_dev(ctx, CH.dt_dev(ctx, V), V, torch.empty_like(V), P, torch.empty_like(P))
    during = _free()            # everything above is allocated now
    del cg
    ctx.close()
    del U, Un, G, R, V, P
    return during


def test_contexts_return_their_device_memory():
    from xyst_amd import chorin as CH
    from xyst_amd import riemann as X
    from xyst_amd.cg import CSR
    from xyst_amd.meshgen import box_mesh
    m = box_mesh(32, seed=2)
    vol = X.mesh_vol(m["coord"], m["inpoel"])
    edge, dint = X.mesh_domint(m["coord"], m["inpoel"])
    sup, ints = X.edges_as_superedges(edge, dint)
    sn, sv, besym = X.mesh_symbc(m["coord"], m["triinpoel"], m["faceset"], [1, 2, 3, 4, 5, 6])
    ce, cd = CH.mesh_domint(m["coord"], m["inpoel"])
    chsup, chints = X.edges_as_superedges(ce, cd)
    A = CSR.poisson(m["coord"], m["inpoel"])
    _one_lifetime(m, vol, sn, sv, besym, sup, ints, chsup, chints, A, 1)   # warm-up
    f0 = _free()
    for k in range(4):
        during = _one_lifetime(m, vol, sn, sv, besym, sup, ints, chsup, chints, A, 3)
        f = _free()
        _say("  lifetime %d: %.1f MiB in use during it, free memory change after %.1f MiB"
             % (k, (f0 - during) / 2 ** 20, (f - f0) / 2 ** 20))
        # the measurement sees a lifetime's allocations ...
        assert f0 - during > 32 * 2 ** 20
    # ... and none of them is left once its context is gone
    assert f0 - f < 4 * 2 ** 20
