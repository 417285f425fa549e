"""ExodusII mesh input and nodal-field output through libxg's NetCDF-classic
reader/writer (xg_exo_* in include/xg.h): the file side of the path, in
place of IO/ExodusIIMeshReader and IO/ExodusIIMeshWriter.

`read_mesh` returns the mesh dictionary the solvers take (the layout
`meshgen.box_mesh` produces): coord (3, N), inpoel (T, 4), triinpoel (F, 3)
and faceset (F,) side-set ids.  `Writer` creates a file with the mesh and
node-variable names and appends steps (writeTimeStamp / writeNodeScalar),
and `read_fields` reads a step back.
"""
import ctypes as C

import numpy as np

from ._xg import XG_OK, lib


class ExodusError(RuntimeError):
    pass


class _Handle:
    def __init__(self, path, append=False, create_args=None):
        L = lib()
        h = C.c_void_p()
        if create_args is None:
            rc = L.xg_exo_open(str(path).encode(), 1 if append else 0, C.byref(h))
        else:
            rc = L.xg_exo_create(str(path).encode(), *create_args, C.byref(h))
        self.h = h
        if rc != XG_OK:
            msg = L.xg_exo_error(h).decode() if h else "out of memory"
            self.close()
            raise ExodusError(msg)

    def check(self, rc):
        if rc != XG_OK:
            raise ExodusError(lib().xg_exo_error(self.h).decode())

    def close(self):
        if self.h:
            lib().xg_exo_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sizes(self):
        v = [C.c_int64(), C.c_int64(), C.c_int64(), C.c_int(), C.c_int(), C.c_int64()]
        self.check(lib().xg_exo_sizes(self.h, *[C.byref(x) for x in v]))
        return dict(npoin=v[0].value, ntet=v[1].value, ntri=v[2].value,
                    nsideset=v[3].value, nnodevar=v[4].value, nstep=v[5].value)


def _tet_oriented(tri, inpoel):
    """Boundary triangles oriented as the reference's partitioner rebuilds
    them from their tetrahedra: one of {t0,t2,t1}, {t0,t1,t3}, {t0,t3,t2},
    {t1,t2,t3} (Partitioner.cpp:578-587), i.e. outward for positively
    oriented tetrahedra.  Faces are matched by their sorted node triples."""
    inpoel = np.asarray(inpoel, np.int64)
    faces = np.concatenate([inpoel[:, [0, 2, 1]], inpoel[:, [0, 1, 3]],
                            inpoel[:, [0, 3, 2]], inpoel[:, [1, 2, 3]]])

    def keys(f):
        k = np.ascontiguousarray(np.sort(f, axis=1))
        return k.view(np.dtype((np.void, 24))).ravel()

    fk, tk = keys(faces), keys(np.asarray(tri, np.int64))
    order = np.argsort(fk, kind="stable")
    pos = np.searchsorted(fk[order], tk)
    pos = np.minimum(pos, fk.size - 1)
    hit = order[pos]
    bad = fk[hit] != tk
    if bad.any():
        raise ExodusError("side-set triangle %s is not a tetrahedron face"
                          % (tuple(np.sort(tri[np.argmax(bad)])),))
    return faces[hit]


def read_mesh(path, orient=True):
    """Coordinates, tetrahedra and side-set triangles of an ExodusII file.
    With orient=True the triangles are re-oriented from their tetrahedra as
    the reference's partitioner does before the solver sees them."""
    with _Handle(path) as h:
        s = h.sizes()
        N = s["npoin"]
        coord = np.empty((3, N))
        h.check(lib().xg_exo_coords(h.h, coord[0].ctypes.data, coord[1].ctypes.data,
                                    coord[2].ctypes.data))
        inpoel = np.empty((s["ntet"], 4), np.int64)
        if s["ntet"]:
            h.check(lib().xg_exo_tets(h.h, inpoel.ctypes.data))
        tris, sets = [], []
        for k in range(s["nsideset"]):
            sid, nf = C.c_int(), C.c_int64()
            h.check(lib().xg_exo_sideset(h.h, k, C.byref(sid), C.byref(nf)))
            t = np.empty((nf.value, 3), np.int64)
            if nf.value:
                h.check(lib().xg_exo_sideset_faces(h.h, k, t.ctypes.data))
            tris.append(t)
            sets.append(np.full(nf.value, sid.value, np.int32))
    tri = np.concatenate(tris) if tris else np.empty((0, 3), np.int64)
    faceset = np.concatenate(sets) if sets else np.empty(0, np.int32)
    if orient and tri.size:
        tri = _tet_oriented(tri, inpoel)
    return {"coord": coord, "inpoel": inpoel, "triinpoel": tri, "faceset": faceset}


def read_fields(path, step=-1):
    """(names, fields (N, nvar), time) of one step (negative: from the end)."""
    with _Handle(path) as h:
        s = h.sizes()
        ns = s["nstep"]
        if ns == 0:
            raise ExodusError("%s holds no steps" % path)
        if step < 0:
            step += ns
        if not 0 <= step < ns:
            raise ExodusError("step out of range")
        names = []
        buf = C.create_string_buffer(256)
        for v in range(s["nnodevar"]):
            h.check(lib().xg_exo_node_var_name(h.h, v, buf, 256))
            names.append(buf.value.decode())
        fields = np.empty((s["npoin"], s["nnodevar"]))
        col = np.empty(s["npoin"])
        for v in range(s["nnodevar"]):
            h.check(lib().xg_exo_read_node_var(h.h, step, v, col.ctypes.data))
            fields[:, v] = col
        t = np.empty(ns)
        h.check(lib().xg_exo_times(h.h, t.ctypes.data))
    return names, fields, float(t[step])


def times(path):
    with _Handle(path) as h:
        t = np.empty(h.sizes()["nstep"])
        if t.size:
            h.check(lib().xg_exo_times(h.h, t.ctypes.data))
    return t


class Writer:
    """ExodusIIMeshWriter: Writer(path, mesh, names) creates the file (mesh +
    node-variable names); Writer(path, append=True) reopens one it wrote.
    write_step(t, fields) appends a step; fields is (N, nvar)."""

    def __init__(self, path, mesh=None, names=(), append=False):
        if append:
            self._h = _Handle(path, append=True)
        else:
            coord = np.ascontiguousarray(mesh["coord"], dtype=np.float64)
            inpoel = np.ascontiguousarray(mesh["inpoel"], dtype=np.int64)
            tri = np.ascontiguousarray(mesh.get("triinpoel", np.empty((0, 3))),
                                       dtype=np.int64)
            faceset = np.asarray(mesh.get("faceset", np.empty(0)), dtype=np.int64)
            ids = np.unique(faceset).astype(np.int32)
            order = np.argsort(faceset, kind="stable")
            off = np.searchsorted(faceset[order], ids).astype(np.int64)
            off = np.append(off, faceset.size).astype(np.int64)
            face = np.ascontiguousarray(order, dtype=np.int64)
            x, y, z = (np.ascontiguousarray(c) for c in coord)
            enc = [n.encode() for n in names]
            arr = (C.c_char_p * max(1, len(enc)))(*enc)
            self._keep = (x, y, z, inpoel, tri, ids, off, face, arr)
            args = (coord.shape[1], x.ctypes.data, y.ctypes.data, z.ctypes.data,
                    inpoel.shape[0], inpoel.ctypes.data if inpoel.size else None,
                    tri.shape[0], tri.ctypes.data if tri.size else None,
                    int(ids.size), ids.ctypes.data if ids.size else None,
                    off.ctypes.data if ids.size else None,
                    face.ctypes.data if ids.size else None,
                    len(enc), C.cast(arr, C.c_void_p) if enc else None)
            self._h = _Handle(path, create_args=args)
        s = self._h.sizes()
        self.npoin, self.nvar, self.nstep = s["npoin"], s["nnodevar"], s["nstep"]

    def write_step(self, t, fields=None):
        step = self.nstep
        self._h.check(lib().xg_exo_write_time(self._h.h, step, float(t)))
        if fields is not None:
            f = np.asarray(fields, dtype=np.float64).reshape(self.npoin, -1)
            if f.shape[1] != self.nvar:
                raise ExodusError("expected %d node variables" % self.nvar)
            for v in range(self.nvar):
                col = np.ascontiguousarray(f[:, v])
                self._h.check(lib().xg_exo_write_node_var(self._h.h, step, v,
                                                          col.ctypes.data))
        self.nstep += 1
        return step

    def close(self):
        self._h.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
