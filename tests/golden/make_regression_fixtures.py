"""Extract the reference's RieCG/Sod, KozCG/Sod, ChoCG/Lid,
ChoCG/Poiseuille, analytic-problem (riecg_analytic.npz, see
analytic_fixtures), RieCG/SlotCyl, ChoCG/Poisson and ChoCG/SlotCyl
regression fixtures into riecg_sod.npz, kozcg_sod.npz, chocg_lid.npz,
chocg_poiseuille.npz, riecg_analytic.npz, riecg_slotcyl.npz,
chocg_poisson.npz and chocg_slotcyl.npz (the argument "poisson" or
"chocg_slotcyl" writes only that one).

Source files (reference test data, read as data only), per directory
tests/regression/inciter/{RieCG,KozCG}/Sod:
  rectangle_01_1.5k.exo   input mesh (the same file for both)
  sod.std.exo             fields after 10 steps
  diag.std                per-step diagnostics
tests/regression/inciter/ChoCG/Lid: unitcube_1k.exo, ldc.std.exo,
diag_ldc.std, and tests/regression/inciter/ChoCG/Poiseuille:
poiseuille1tetz.exo (side sets on tet sides), poiseuille_{damp2,damp4,rk2,
rk3,rk4}.std.exo and diag_poiseuille_*.std (also _theta).
Run configurations (sod.q, ldc.q):
  RieCG: nstep 10, term 0.2, cfl 0.5, gamma 1.4, symmetry on side sets 2,4,5,6.
  KozCG: nstep 10, cfl 0.5, gamma 1.4, symmetry on 2,4,5,6, Dirichlet (all
         five components, initial values) on side sets 1 and 3, FCT with
         fctclip = true and fctsys = {1, 2, 5}, fctdif at its default 1.0.
  ChoCG: nstep 10, cfl 0.9, flux damp4, Re 100 (viscosity 0.01), rk 1,
         noslip on side sets 1,2,3,5,6, velocity (1,0,0) set on side set 4
         (mask 2), pressure CG: 500 iterations, tol 1e-3, Jacobi, hydrostat
         node 0; fields written every 2 steps (the last is step 10).
  ChoCG Poiseuille (poiseuille_*.q, one PE): nstep 20, viscosity 0.01,
         damp2 / damp4 / damp2 with rk 2, 3, 4 (flux, cfl and rk read from
         each case's .q file); noslip on side sets 3, 4; v = w = initial value (mask 1) on
         sets 1 and 5; pressure Dirichlet p = 2.4 on set 1 and 0 on set 2
         (mask 2), CG tol 1e-3 with Jacobi, 500 iterations; the theta case
         (theta 0.5, damp2, momentum CG tol 1e-3, 50 iterations, Jacobi)
         keeps its diagnostics only (its fields are split over two
         partitions).
The output mesh is renumbered by the solver, so output nodes are matched to
input nodes by coordinates.
"""
import os
import re
import shutil
import sys

import numpy as np
from scipy.io import netcdf_file

REF = "/root/reference/tests/regression/inciter"
HERE = os.path.dirname(os.path.abspath(__file__))


# ExodusII TETRA4 sides (1-based side number -> local nodes)
TET_SIDES = {1: (0, 1, 3), 2: (1, 2, 3), 3: (0, 3, 2), 4: (0, 2, 1)}


def read_mesh(ref, name="rectangle_01_1.5k.exo"):
    """Coordinates, tets and boundary triangles with their side-set ids.  Side
    sets may name the sides of a triangle block (element block 1 of
    triangles, tets in block 2) or tet sides (a single tet block)."""
    f = netcdf_file(os.path.join(ref, name), "r", mmap=False)
    coord = np.stack([f.variables[c].data.astype(np.float64)
                      for c in ("coordx", "coordy", "coordz")])
    blk1 = f.variables["connect1"].data.astype(np.int64) - 1
    if blk1.shape[1] == 3:
        tri_blk, tets = blk1, f.variables["connect2"].data.astype(np.int64) - 1
    else:
        tri_blk, tets = None, blk1
    tris, sets = [], []
    nss = f.dimensions["num_side_sets"]
    ids = f.variables["ss_prop1"].data
    for s in range(1, nss + 1):
        elem = f.variables["elem_ss%d" % s].data.astype(np.int64)
        if tri_blk is not None:
            assert (elem <= tri_blk.shape[0]).all(), "side set on tet sides"
            tris.append(tri_blk[elem - 1])
        else:
            side = f.variables["side_ss%d" % s].data.astype(np.int64)
            tris.append(np.array([tets[e - 1][list(TET_SIDES[k])]
                                  for e, k in zip(elem, side)], np.int64))
        sets.append(np.full(elem.size, ids[s - 1], np.int32))
    tri = np.concatenate(tris)
    faceset = np.concatenate(sets)
    # The solver does not use the file's triangle orientation: the
    # partitioner rebuilds each boundary face from its tet as one of
    # {t0,t2,t1}, {t0,t1,t3}, {t0,t3,t2}, {t1,t2,t3} (Partitioner.cpp:578-587),
    # i.e. outward for positively oriented tets.
    tetface = {}
    for t in tets:
        for fc in ((t[0], t[2], t[1]), (t[0], t[1], t[3]), (t[0], t[3], t[2]),
                   (t[1], t[2], t[3])):
            tetface[tuple(sorted(fc))] = fc
    tri = np.array([tetface[tuple(sorted(fc))] for fc in tri], np.int64)
    return coord, tets, tri, faceset


def read_result(ref, coord, name="sod.std.exo"):
    g = netcdf_file(os.path.join(ref, name), "r", mmap=False)
    Y = np.stack([g.variables[c].data.astype(np.float64)
                  for c in ("coordx", "coordy", "coordz")])
    key = lambda P: np.lexsort(np.round(P, 12)[::-1])
    ia, ib = key(coord), key(Y)
    assert np.abs(coord[:, ia] - Y[:, ib]).max() < 1e-12
    perm = np.empty(coord.shape[1], np.int64)
    perm[ia] = ib      # input node i -> output node perm[i]
    names = [b"".join(x).decode().strip("\x00 ")
             for x in g.variables["name_nod_var"].data]
    fields = np.stack([g.variables["vals_nod_var%d" % (i + 1)].data[-1][perm]
                       for i in range(len(names))], 1)
    t_final = float(g.variables["time_whole"].data[-1])
    return fields, names, t_final


def main():
    for solver, out, extra in (
            ("RieCG", "riecg_sod.npz", dict(term=0.2)),
            ("KozCG", "kozcg_sod.npz",
             dict(term=1.0e300, dirsets=np.array([1, 3], np.int32), fctclip=1,
                  fctsys=np.array([0, 1, 4], np.int32), fctdif=1.0))):
        ref = os.path.join(REF, solver, "Sod")
        coord, tets, tri, faceset = read_mesh(ref)
        fields, names, t_final = read_result(ref, coord)
        diag = np.loadtxt(os.path.join(ref, "diag.std"))
        np.savez_compressed(os.path.join(HERE, out), coord=coord, inpoel=tets,
                            triinpoel=tri, faceset=faceset, fields=fields,
                            field_names=np.array(names), t_final=t_final,
                            diag=diag, symsets=np.array([2, 4, 5, 6], np.int32),
                            nstep=10, cfl=0.5, gamma=1.4, **extra)
        print("wrote", out, fields.shape, names, t_final)
    ref = os.path.join(REF, "ChoCG", "Lid")
    coord, tets, tri, faceset = read_mesh(ref, "unitcube_1k.exo")
    fields, names, t_final = read_result(ref, coord, "ldc.std.exo")
    diag = np.loadtxt(os.path.join(ref, "diag_ldc.std"))
    np.savez_compressed(os.path.join(HERE, "chocg_lid.npz"), coord=coord, inpoel=tets,
                        triinpoel=tri, faceset=faceset, fields=fields,
                        field_names=np.array(names), t_final=t_final, diag=diag,
                        nstep=10, cfl=0.9, mu=0.01, flux="damp4", rk=1,
                        noslip=np.array([1, 2, 3, 5, 6], np.int32), dirset=4,
                        dirmask=np.array([2, 2, 2], np.int32),
                        dirval=np.array([1.0, 0.0, 0.0]), p_iter=500, p_tol=1e-3,
                        p_hydrostat=0)
    print("wrote chocg_lid.npz", fields.shape, names, t_final)
    ref = os.path.join(REF, "ChoCG", "Poiseuille")
    coord, tets, tri, faceset = read_mesh(ref, "poiseuille1tetz.exo")
    cases = {}
    for case in ("damp2", "damp4", "rk2", "rk3", "rk4", "theta"):
        q = open(os.path.join(ref, "poiseuille_%s.q" % case)).read()
        cfl = float(re.search(r"^cfl\s*=\s*([0-9.eE+-]+)", q, re.M).group(1))
        flux = re.search(r'^flux\s*=\s*"(\w+)"', q, re.M).group(1)
        m = re.search(r"^rk\s*=\s*(\d+)", q, re.M)
        rk = int(m.group(1)) if m else 1
        m = re.search(r"^theta\s*=\s*([0-9.eE+-]+)", q, re.M)
        theta = float(m.group(1)) if m else 0.0
        diag = np.loadtxt(os.path.join(ref, "diag_poiseuille_%s.std" % case))
        cases.update({case + "_diag": diag, case + "_flux": flux, case + "_cfl": cfl,
                      case + "_rk": rk, case + "_theta": theta})
        if case != "theta":          # two-partition output files for theta
            fields, names, t_final = read_result(ref, coord, "poiseuille_%s.std.exo" % case)
            cases.update({case + "_fields": fields, case + "_t_final": t_final})
    np.savez_compressed(os.path.join(HERE, "chocg_poiseuille.npz"), coord=coord,
                        inpoel=tets, triinpoel=tri, faceset=faceset,
                        field_names=np.array(names), nstep=20, mu=0.01,
                        noslip=np.array([3, 4], np.int32),
                        dirsets=np.array([1, 5], np.int32),
                        dirmask=np.array([0, 1, 1], np.int32),
                        p_dirsets=np.array([1, 2], np.int32),
                        p_dirvals=np.array([2.4, 0.0]), p_iter=500, p_tol=1e-3,
                        **cases)
    print("wrote chocg_poiseuille.npz", coord.shape, tets.shape, tri.shape)

    analytic_fixtures()
    slotcyl_fixtures()
    poisson_fixtures()
    chocg_slotcyl_fixtures()

    # ExodusII files kept as they are, for the file reader's tests: a mesh
    # with a triangle block and side sets on it (RieCG/Sod), a mesh with
    # untyped TRI/TETRA blocks (ChoCG/Lid), result files with node sets, node
    # variables and time steps (ChoCG/Lid and RieCG/Sod output; the latter is
    # also the end-to-end RieCG test's reference)
    dst = os.path.join(HERE, "exo")
    os.makedirs(dst, exist_ok=True)
    for sub, name in (("RieCG/Sod", "rectangle_01_1.5k.exo"),
                      ("ChoCG/Lid", "unitcube_1k.exo"), ("ChoCG/Lid", "ldc.std.exo"),
                      ("RieCG/Sod", "sod.std.exo")):
        shutil.copyfile(os.path.join(REF, sub, name), os.path.join(dst, name))
    print("copied 4 ExodusII files to", dst)


def merge_partitions(ref, coord, stem, nfile):
    """Fields of a partitioned result (stem.0 .. stem.<nfile-1>) matched to
    the input nodes by coordinates; a node shared by partitions takes the
    value of the first file that holds it."""
    key = lambda P: [tuple(np.round(P[:, i], 12)) for i in range(P.shape[1])]
    index = {k: i for i, k in enumerate(key(coord))}
    fields = None
    seen = np.zeros(coord.shape[1], bool)
    for k in range(nfile):
        g = netcdf_file(os.path.join(ref, "%s.%d" % (stem, k)), "r", mmap=False)
        Y = np.stack([g.variables[c].data.astype(np.float64)
                      for c in ("coordx", "coordy", "coordz")])
        names = [b"".join(x).decode().strip("\x00 ")
                 for x in g.variables["name_nod_var"].data]
        if fields is None:
            fields = np.zeros((coord.shape[1], len(names)))
        for j, kk in enumerate(key(Y)):
            i = index[kk]
            if not seen[i]:
                seen[i] = True
                fields[i] = [g.variables["vals_nod_var%d" % (v + 1)].data[-1][j]
                             for v in range(len(names))]
        t_final = float(g.variables["time_whole"].data[-1])
    assert seen.all()
    return fields, names, t_final


def chare_of_tet(ref, coord, tets, stem, nfile):
    """Partition k of each input tet: the file stem.k whose connectivity holds
    the tet (matched by the coordinates of its corners)."""
    key = lambda P: [tuple(np.round(P[:, i], 12)) for i in range(P.shape[1])]
    index = {k: i for i, k in enumerate(key(coord))}
    tet_id = {tuple(sorted(t)): e for e, t in enumerate(tets)}
    owner = np.full(tets.shape[0], -1, np.int8)
    for k in range(nfile):
        g = netcdf_file(os.path.join(ref, "%s.%d" % (stem, k)), "r", mmap=False)
        Y = np.stack([g.variables[c].data.astype(np.float64)
                      for c in ("coordx", "coordy", "coordz")])
        glob = np.array([index[kk] for kk in key(Y)])
        for t in g.variables["connect1"].data.astype(np.int64) - 1:
            e = tet_id[tuple(sorted(glob[t]))]
            assert owner[e] == -1
            owner[e] = k
    assert (owner >= 0).all()
    return owner


def analytic_fixtures():
    """RieCG/VorticalFlow (rusanov, stab2, hllc, hllc+stab2, steady),
    RieCG/TaylorGreen, KozCG/VorticalFlow, KozCG/TaylorGreen (unitcube_1k,
    Dirichlet on all six side sets, every component set to the initial
    condition) and RieCG/Pipe (rectangle_01_1.5k, symmetry on 2,4,5,6,
    pressure BCs on sets 1 and 3) into riecg_analytic.npz.
    Run configurations (the .q files): vortical_flow alpha 0.1, kappa 1.0,
    p0 10.0, gamma 5/3, cfl 0.8, term 1.0 (steady: nstep 10, local time
    steps); taylor_green gamma 5/3, cfl 0.8, term 1.0, diagnostics every
    second step; KozCG cases without FCT; pipe gamma 1.4, cfl 0.5, nstep 20,
    IC rho 1, u 0, p 99950, bc_pre (1.0, 100000) on set 1 and (1.0, 99900) on
    set 3."""
    out = {}
    rv = os.path.join(REF, "RieCG", "VorticalFlow")
    coord, tets, tri, faceset = read_mesh(rv, "unitcube_1k.exo")
    out.update(coord=coord, inpoel=tets, triinpoel=tri, faceset=faceset)
    for case, diagf, exo in (("rusanov", "diag.std", "vortical_flow.std.exo"),
                             ("stab2", "diag_stab2.std", "vortical_flow_stab2.std.exo"),
                             ("hllc", "diag_hllc.std", None),
                             ("hllc_stab2", "diag_hllc_stab2.std", None),
                             ("steady", "diag_steady.std", "vortical_flow_steady.std.exo")):
        out["vf_%s_diag" % case] = np.loadtxt(os.path.join(rv, diagf))
        if exo:
            f, names, tf = read_result(rv, coord, exo)
            out["vf_%s_fields" % case] = f[:, :6]
            out["vf_%s_t_final" % case] = tf
    for case, stem in (("hllc", "vortical_flow_hllc_pe4_u0.5.std.exo"),
                       ("hllc_stab2", "vortical_flow_hllc_stab2_pe4_u0.5.std.exo")):
        f, names, tf = merge_partitions(rv, coord, stem, 8)
        out["vf_%s_fields" % case] = f[:, :6]
        out["vf_%s_t_final" % case] = tf
    # the HLLC runs are partitioned (4 PEs, 8 chares): the chare of every
    # tet, from the tets each partition file holds
    out["vf_chare_of_tet"] = chare_of_tet(rv, coord, tets,
                                          "vortical_flow_hllc_pe4_u0.5.std.exo", 8)
    rt = os.path.join(REF, "RieCG", "TaylorGreen")
    out["tg_diag"] = np.loadtxt(os.path.join(rt, "diag.std"))
    f, names, tf = read_result(rt, coord, "taylor_green.std.exo")
    out.update(tg_fields=f[:, :6], tg_t_final=tf)
    kv = os.path.join(REF, "KozCG", "VorticalFlow")
    out["koz_vf_diag"] = np.loadtxt(os.path.join(kv, "diag.std"))
    f, names, tf = read_result(kv, coord, "vortical_flow.std.exo")
    out.update(koz_vf_fields=f[:, :6], koz_vf_t_final=tf)
    kt = os.path.join(REF, "KozCG", "TaylorGreen")
    out["koz_tg_diag"] = np.loadtxt(os.path.join(kt, "diag.std"))
    f, names, tf = read_result(kt, coord, "taylor_green.std.exo")
    out.update(koz_tg_fields=f[:, :6], koz_tg_t_final=tf)
    rp = os.path.join(REF, "RieCG", "Pipe")
    pc, pt, ptri, pfs = read_mesh(rp)
    f, names, tf = read_result(rp, pc, "pipe.std.exo")
    out.update(pipe_coord=pc, pipe_inpoel=pt, pipe_triinpoel=ptri, pipe_faceset=pfs,
               pipe_fields=f, pipe_t_final=tf,
               pipe_diag=np.loadtxt(os.path.join(rp, "diag.std")))
    out["field_names"] = np.array(["density", "xvelocity", "yvelocity", "zvelocity",
                                   "energy", "pressure"])
    np.savez_compressed(os.path.join(HERE, "riecg_analytic.npz"), **out)
    print("wrote riecg_analytic.npz", sorted(out))


def slotcyl_fixtures():
    """RieCG/SlotCyl (one transported scalar c0: slotted cylinder, cone and
    hump rotated by a prescribed flow; Rusanov and HLLC runs, both compared
    with the same diag.std) into riecg_slotcyl.npz.  Run configuration
    (slot_cyl.q, slot_cyl_hllc.q): unitsquare_01_3.6k.exo, nstep 20, cfl
    0.9, gamma 5/3, Dirichlet to the initial condition on side set 1 (all
    six components) and side set 2 (the five flow components, c0 free);
    fields (incl. the analytic ones) written after the last step; the
    Rusanov run also as 8 chares (slot_cyl_pe4_u0.5.std.exo.*); and KozCG/SlotCyl
    (koz_*: cfl 0.5, FCT defaults, freezeflow 3 from time 0)."""
    rs = os.path.join(REF, "RieCG", "SlotCyl")
    coord, tets, tri, faceset = read_mesh(rs, "unitsquare_01_3.6k.exo")
    out = dict(coord=coord, inpoel=tets, triinpoel=tri, faceset=faceset,
               diag=np.loadtxt(os.path.join(rs, "diag.std")), nstep=20, cfl=0.9,
               gamma=5.0 / 3.0)
    for case, exo in (("rusanov", "slot_cyl.std.exo"), ("hllc", "slot_cyl_hllc.std.exo")):
        f, names, tf = read_result(rs, coord, exo)
        out["%s_fields" % case] = f
        out["%s_t_final" % case] = tf
    out["field_names"] = np.array(names)
    # the Rusanov run on 4 PEs with 8 chares (-u 0.5): fields and the chare
    # of every tet, from the partition files
    f, _, tf = merge_partitions(rs, coord, "slot_cyl_pe4_u0.5.std.exo", 8)
    out.update(pe8_fields=f, pe8_t_final=tf,
               chare_of_tet=chare_of_tet(rs, coord, tets, "slot_cyl_pe4_u0.5.std.exo", 8))
    # KozCG/SlotCyl (serial): the same mesh and BCs, cfl 0.5, FCT with the
    # defaults, freezeflow 3 from freezetime 0
    ks = os.path.join(REF, "KozCG", "SlotCyl")
    f, _, tf = read_result(ks, coord, "slot_cyl.std.exo")
    out.update(koz_fields=f, koz_t_final=tf, koz_diag=np.loadtxt(os.path.join(ks, "diag.std")),
               koz_cfl=0.5, koz_freezeflow=3.0, koz_freezetime=0.0)
    np.savez_compressed(os.path.join(HERE, "riecg_slotcyl.npz"), **out)
    print("wrote riecg_slotcyl.npz", sorted(out), names)


def poisson_fixtures():
    """ChoCG/Poisson (one pressure solve, nstep 1, dt 0.5; CG tol 1e-6, 100
    iterations, no preconditioner unless noted) into chocg_poisson.npz.
    Run configurations (poisson_*.q): poisson_const, poisson_sine and
    poisson_sine3 on unitcube_01_1k.exo with pressure Dirichlet (mask 1: the
    problem's pressure) on side sets 1-6; poisson_neumann on
    unitcube_0pidiv4_1k.exo with Dirichlet on sets 1, 2, 3 and Neumann
    (bc_sym) on 4, 5, 6; poisson_sine3 also with the Jacobi preconditioner
    on 4 PEs, compared with the same diag file.  Fields (pressure, analytic)
    from the serial run where the reference holds one (poisson_const) and
    from the partition files otherwise."""
    ref = os.path.join(REF, "ChoCG", "Poisson")
    out = {}
    for key, mesh in (("cube", "unitcube_01_1k.exo"), ("cube4", "unitcube_0pidiv4_1k.exo")):
        coord, tets, tri, faceset = read_mesh(ref, mesh)
        out.update({key + "_coord": coord, key + "_inpoel": tets, key + "_triinpoel": tri,
                    key + "_faceset": faceset})
    for case, mesh, files, ndiag in (
            ("const", "cube", ("poisson_const.std.exo", 0), "const"),
            ("sine", "cube", ("poisson_sine_pe2.std.exo", 2), "sine"),
            ("sine3", "cube", ("poisson_sine3_pe2.std.exo", 2), "sine3"),
            ("sine3_jacobi", "cube", ("poisson_sine3_jacobi_pe4.std.exo", 4), "sine3"),
            ("neumann", "cube4", ("poisson_neumann_pe2.std.exo", 2), "neumann")):
        coord = out[mesh + "_coord"]
        if files[1] == 0:
            fields, names, _ = read_result(ref, coord, files[0])
        else:
            fields, names, _ = merge_partitions(ref, coord, files[0], files[1])
        out[case + "_fields"] = fields
        out[case + "_names"] = np.array(names)
        out[case + "_diag"] = np.loadtxt(os.path.join(ref, "diag_poisson_%s.std" % ndiag))
    out.update(p_tol=1e-6, p_iter=100, dt=0.5)
    np.savez_compressed(os.path.join(HERE, "chocg_poisson.npz"), **out)
    print("wrote chocg_poisson.npz", sorted(out))


def chocg_slotcyl_fixtures():
    """ChoCG/SlotCyl (velocity and one transported scalar c0: the slotted
    cylinder, cone and hump carried by a prescribed rotation held by a
    momentum source) into chocg_slotcyl.npz.  Run configuration
    (slot_cyl*.q): unitsquare_01_3.6k.exo, nstep 20, cfl 0.9, Dirichlet to
    the initial condition on side set 1 (all four components) and side set 2
    (velocity only), pressure CG with Jacobi, tol 1e-2, 300 iterations,
    hydrostat node 0; damp2 with rk 3 (slot_cyl.q), damp4 with rk 4
    (slot_cyl_damp4.q) and damp4 with rk 4, freezeflow 2 after t = 0.1
    (slot_cyl_damp4_freeze.q); fields after the last step."""
    ref = os.path.join(REF, "ChoCG", "SlotCyl")
    coord, tets, tri, faceset = read_mesh(ref, "unitsquare_01_3.6k.exo")
    out = dict(coord=coord, inpoel=tets, triinpoel=tri, faceset=faceset, nstep=20, cfl=0.9,
               p_iter=300, p_tol=1e-2, p_hydrostat=0, freezeflow=2.0, freezetime=0.1)
    for case, exo, diag in (("damp2", "slot_cyl.std.exo", "diag.std"),
                            ("damp4", "slot_cyl_damp4.std.exo", "diag_damp4.std"),
                            ("freeze", "slot_cyl_damp4_freeze.std.exo",
                             "diag_damp4_freeze.std")):
        f, names, tf = read_result(ref, coord, exo)
        out.update({case + "_fields": f, case + "_t_final": tf,
                    case + "_diag": np.loadtxt(os.path.join(ref, diag))})
    out["field_names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "chocg_slotcyl.npz"), **out)
    print("wrote chocg_slotcyl.npz", names)


if __name__ == "__main__":
    if sys.argv[1:] == ["poisson"]:
        sys.exit(poisson_fixtures())
    if sys.argv[1:] == ["chocg_slotcyl"]:
        sys.exit(chocg_slotcyl_fixtures())
    sys.exit(main())
