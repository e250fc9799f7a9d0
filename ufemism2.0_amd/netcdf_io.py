"""Mesh and restart files in the reference's NetCDF layout.

Meshes (setup_mesh_in_netcdf_file, netcdf_setup_grid_mesh_in_file.f90:130-453;
setup_mesh_from_file, netcdf_setup_grid_mesh_from_file.f90:133-241): the
dimensions vi, ti, ci (nC_mem), two, three; the scalars xmin, xmax, ymin,
ymax, tol_dist, lambda_M, phi_M, beta_stereo; the primary arrays V, nC, C,
niTri, iTri, VBI, Tri, Tricc, TriC (field names of
netcdf_field_name_options.f90:41-81), 1-based with 0 where a row is shorter
than ci or a triangle has no neighbour, and the secondary ones this module
also writes for convenience (TriBI, TriGC, A, R).  A reader, like the
reference's, takes only the primary arrays and recomputes the rest.

Velocity restarts (create_restart_file_SSA / write_to_restart_file_SSA,
SSA_main.f90:538-627; the DIVA pair, DIVA_main.f90:808-910): the mesh, an
unlimited time dimension with its variable, a zeta dimension for 3-D fields,
and per time slice u_b, v_b (SSA) or u_vav_b, v_vav_b, tau_bx_b, tau_by_b,
eta_3D_b, u_base_b, v_base_b (DIVA).

Dimension order.  The reference writes through the Fortran NetCDF interface,
which lists dimensions fastest first; a file therefore stores V(vi, two) as a
(two, vi) array in C order, u_b(ti, time) as (time, ti) and eta_3D_b(ti, zeta,
time) as (time, zeta, ti).  This module reads and writes that C order and
hands out arrays in the engine's layout ((nV, 2), (nT,), (nz*nT,) with the
triangle fastest, as the reference's (nT, nz) column-major array).

Formats.  Writing produces NetCDF classic files with 64-bit offsets (CDF-2,
written here directly; scipy's writer misplaces scalar variables after the
record section), which the reference's netCDF library opens like its own
NetCDF-4 output.  Reading accepts classic files and NetCDF-4 files, which are
HDF5 files with one dataset per variable: those are read through the HDF5 C
library (ctypes; libhdf5 is looked up with ctypes.util.find_library, then in
$UFE_HDF5_LIB, then under the Python prefix); without it a NetCDF-4 file
raises an error that says so.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import os
import sys

import numpy as np

from .mesh import Mesh, vertex_geometry

MESH_SCALARS = ("xmin", "xmax", "ymin", "ymax", "tol_dist", "lambda_M", "phi_M", "beta_stereo")
SSA_RESTART_FIELDS = ("u_b", "v_b")
DIVA_RESTART_FIELDS = ("u_vav_b", "v_vav_b", "tau_bx_b", "tau_by_b", "eta_3D_b", "u_base_b", "v_base_b")
FIELD_UNITS = {"u_b": "m/yr", "v_b": "m/yr", "u_vav_b": "m/yr", "v_vav_b": "m/yr", "tau_bx_b": "Pa",
               "tau_by_b": "Pa", "eta_3D_b": "Pa yr", "u_base_b": "m/yr", "v_base_b": "m/yr"}


class MeshFileError(RuntimeError):
    pass


# --------------------------------------------------------------------------
# readers: classic (scipy) and NetCDF-4 (HDF5 through ctypes)
# --------------------------------------------------------------------------

class _ClassicFile:
    def __init__(self, path):
        from scipy.io import netcdf_file
        self.f = netcdf_file(path, "r", mmap=False)

    def has(self, name):
        return name in self.f.variables

    def read(self, name):
        v = self.f.variables[name]
        a = np.array(v.data, copy=True)
        return a.reshape(()) if v.shape == () else a

    def dims(self):
        return dict(self.f.dimensions)

    def close(self):
        self.f.close()


def _find_hdf5():
    cands = [ctypes.util.find_library("hdf5"), os.environ.get("UFE_HDF5_LIB")]
    for pre in (sys.prefix, sys.base_prefix, "/opt/conda"):
        cands.append(os.path.join(pre, "lib", "libhdf5.so"))
    for c in cands:
        if not c:
            continue
        try:
            return C.CDLL(c)
        except OSError:
            continue
    return None


class _Hdf5:
    """The few HDF5 calls a NetCDF-4 reader needs (hid_t is 64-bit since
    HDF5 1.10)."""
    _lib = None

    def __init__(self):
        if _Hdf5._lib is None:
            h = _find_hdf5()
            if h is None:
                raise MeshFileError("this is a NetCDF-4 (HDF5) file and no HDF5 library was found; "
                                    "set UFE_HDF5_LIB to libhdf5.so or convert the file to NetCDF classic")
            hid = C.c_int64
            h.H5open()
            for fn, res, args in (
                    ("H5Fopen", hid, [C.c_char_p, C.c_uint, hid]),
                    ("H5Fcreate", hid, [C.c_char_p, C.c_uint, hid, hid]),
                    ("H5Fclose", C.c_int, [hid]),
                    ("H5Lexists", C.c_int, [hid, C.c_char_p, hid]),
                    ("H5Dopen2", hid, [hid, C.c_char_p, hid]),
                    ("H5Dcreate2", hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
                    ("H5Dget_space", hid, [hid]),
                    ("H5Dget_type", hid, [hid]),
                    ("H5Dread", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                    ("H5Dwrite", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                    ("H5Dclose", C.c_int, [hid]),
                    ("H5Screate_simple", hid, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
                    ("H5Screate", hid, [C.c_int]),
                    ("H5Sget_simple_extent_ndims", C.c_int, [hid]),
                    ("H5Sget_simple_extent_dims", C.c_int, [hid, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
                    ("H5Sclose", C.c_int, [hid]),
                    ("H5Tget_class", C.c_int, [hid]),
                    ("H5Tclose", C.c_int, [hid])):
                f = getattr(h, fn)
                f.restype, f.argtypes = res, args
            _Hdf5._lib = h
        self.h = _Hdf5._lib
        self.double = C.c_int64.in_dll(self.h, "H5T_NATIVE_DOUBLE_g").value
        self.int32 = C.c_int64.in_dll(self.h, "H5T_NATIVE_INT_g").value


class _Netcdf4File:
    H5T_INTEGER = 0

    def __init__(self, path):
        self.x = _Hdf5()
        self.fid = self.x.h.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.fid < 0:
            raise MeshFileError(f"HDF5 could not open {path}")

    def has(self, name):
        return self.x.h.H5Lexists(self.fid, name.encode(), 0) > 0

    def read(self, name):
        h = self.x.h
        d = h.H5Dopen2(self.fid, name.encode(), 0)
        if d < 0:
            raise MeshFileError(f"no variable {name}")
        try:
            sp = h.H5Dget_space(d)
            nd = h.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_uint64 * max(nd, 1))()
            h.H5Sget_simple_extent_dims(sp, dims, None)
            h.H5Sclose(sp)
            shape = tuple(int(dims[i]) for i in range(nd))
            t = h.H5Dget_type(d)
            is_int = h.H5Tget_class(t) == self.H5T_INTEGER
            h.H5Tclose(t)
            out = np.empty(shape, np.int32 if is_int else np.float64)
            if h.H5Dread(d, self.x.int32 if is_int else self.x.double, 0, 0, 0,
                         out.ctypes.data_as(C.c_void_p)) < 0:
                raise MeshFileError(f"HDF5 could not read {name}")
            return out
        finally:
            h.H5Dclose(d)

    def close(self):
        self.x.h.H5Fclose(self.fid)


def _open(path):
    with open(path, "rb") as fh:
        magic = fh.read(8)
    if magic[:3] == b"CDF":
        return _ClassicFile(path)
    if magic == b"\x89HDF\r\n\x1a\n":
        return _Netcdf4File(path)
    raise MeshFileError(f"{path} is neither a NetCDF classic nor a NetCDF-4/HDF5 file")


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------

def _ragged_to_padded(ptr, ind, width):
    n = len(ptr) - 1
    out = np.zeros((n, width), np.int32)
    cnt = np.diff(ptr)
    if cnt.max(initial=0) > width:
        raise MeshFileError(f"a vertex has {cnt.max()} connections, more than ci = {width}")
    rows = np.repeat(np.arange(n), cnt)
    cols = np.arange(len(ind)) - np.repeat(ptr[:-1], cnt)
    out[rows, cols] = ind + 1
    return out


def _padded_to_ragged(cnt, arr):
    cnt = np.asarray(cnt, np.int64)
    if (cnt < 0).any() or (cnt > arr.shape[1]).any():
        raise MeshFileError("connection counts out of range")
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    mask = np.arange(arr.shape[1])[None, :] < cnt[:, None]
    return ptr, (arr[mask] - 1).astype(np.int32)


def _mesh_from_primary(xmin, xmax, ymin, ymax, V, nC, Cpad, niTri, iTripad, VBI, Tri, TriC):
    """Mesh from the reference's primary arrays (0-based already where
    stated), with the secondary data the engine needs: TriBI and TriGC
    (calc_all_secondary_mesh_data)."""
    from .mesh import build_topology
    nV, nT = V.shape[0], Tri.shape[0]
    if Tri.min() < 0 or Tri.max() >= nV:
        raise MeshFileError("Tri refers to vertices outside 1..nV")
    ref = build_topology(V, Tri, TriC, xmin, xmax, ymin, ymax)
    C_ptr, C_ind = _padded_to_ragged(nC, Cpad)
    iTri_ptr, iTri_ind = _padded_to_ragged(niTri, iTripad)
    if C_ind.size and (C_ind.min() < 0 or C_ind.max() >= nV):
        raise MeshFileError("C refers to vertices outside 1..nV")
    if iTri_ind.size and (iTri_ind.min() < 0 or iTri_ind.max() >= nT):
        raise MeshFileError("iTri refers to triangles outside 1..nTri")
    # the file's connection lists must describe the same graph as Tri
    # (test_mesh_is_self_consistent)
    for name, (p, i), (pr, ir) in (("C", (C_ptr, C_ind), (ref.C_ptr, ref.C_ind)),
                                   ("iTri", (iTri_ptr, iTri_ind), (ref.iTri_ptr, ref.iTri_ind))):
        if not np.array_equal(np.diff(p), np.diff(pr)):
            raise MeshFileError(f"{name} is not consistent with Tri")
        src = np.repeat(np.arange(nV), np.diff(p))
        a = np.lexsort((i, src))
        b = np.lexsort((ir, src))
        if not np.array_equal(i[a], ir[b]):
            raise MeshFileError(f"{name} is not consistent with Tri")
    if not np.array_equal(VBI, ref.VBI):
        raise MeshFileError("VBI is not consistent with the vertex coordinates and the domain border")
    return Mesh(float(xmin), float(xmax), float(ymin), float(ymax), ref.V, ref.Tri,
                np.ascontiguousarray(TriC, np.int32), np.ascontiguousarray(VBI, np.int32), ref.TriBI, ref.TriGC,
                iTri_ptr, iTri_ind, C_ptr, C_ind)


def read_mesh(path) -> Mesh:
    """A mesh from a NetCDF mesh, output or restart file of the reference
    (or one written by write_mesh)."""
    f = _open(path)
    try:
        for name in ("V", "nC", "C", "niTri", "iTri", "VBI", "Tri", "TriC") + MESH_SCALARS[:4]:
            if not f.has(name):
                raise MeshFileError(f"no valid {name} variable in {path}")
        sc = {k: float(np.asarray(f.read(k)).reshape(-1)[0]) for k in MESH_SCALARS[:4]}
        V = np.ascontiguousarray(f.read("V").T, np.float64)
        Tri = np.ascontiguousarray(f.read("Tri").T, np.int32) - 1
        TriC = np.ascontiguousarray(f.read("TriC").T, np.int32) - 1
        nC = f.read("nC").astype(np.int32)
        Cpad = np.ascontiguousarray(f.read("C").T, np.int32)
        niTri = f.read("niTri").astype(np.int32)
        iTripad = np.ascontiguousarray(f.read("iTri").T, np.int32)
        VBI = f.read("VBI").astype(np.int32)
    finally:
        f.close()
    if V.ndim != 2 or V.shape[1] != 2 or Tri.shape[1:] != (3,) or TriC.shape != Tri.shape:
        raise MeshFileError("mesh arrays have the wrong shapes")
    return _mesh_from_primary(sc["xmin"], sc["xmax"], sc["ymin"], sc["ymax"], V, nC, Cpad, niTri, iTripad, VBI,
                              Tri, np.maximum(TriC, -1))


class _Var:
    """A variable to be written: name, dimension names, type ('d' or 'i'),
    attributes and data (records first for record variables)."""

    def __init__(self, name, dims, typ, data, **attrs):
        self.name, self.dims, self.typ, self.attrs = name, tuple(dims), typ, attrs
        self.data = np.asarray(data, np.float64 if typ == "d" else np.int32)


_NC_CHAR, _NC_INT, _NC_DOUBLE = 2, 4, 6


def _write_classic(path, dims, variables, gatts=None):
    """Write a NetCDF classic file with 64-bit offsets (CDF-2: big-endian
    header of dimensions, attributes and variables, then the non-record
    variables and the interleaved records).  dims: [(name, length or None
    for the unlimited one)]."""
    import struct
    dim_ix = {n: i for i, (n, _) in enumerate(dims)}
    rec_dim = [n for n, l in dims if l is None]
    nrec = 0
    for v in variables:
        if v.dims and v.dims[0] in rec_dim:
            nrec = max(nrec, v.data.shape[0])

    def pad4(b):
        return b + b"\0" * (-len(b) % 4)

    def name(n):
        e = n.encode()
        return struct.pack(">i", len(e)) + pad4(e)

    def atts(d):
        if not d:
            return struct.pack(">ii", 0, 0)
        out = struct.pack(">ii", 0x0C, len(d))
        for k, val in d.items():
            if isinstance(val, str):
                e = val.encode()
                out += name(k) + struct.pack(">ii", _NC_CHAR, len(e)) + pad4(e)
            else:
                a = np.atleast_1d(np.asarray(val, np.float64))
                out += name(k) + struct.pack(">ii", _NC_DOUBLE, a.size) + a.astype(">f8").tobytes()
        return out

    def is_rec(v):
        return bool(v.dims) and v.dims[0] in rec_dim

    def slab(v):
        """Bytes of the variable (non-record) or of one record slice."""
        n = 1
        for d in (v.dims[1:] if is_rec(v) else v.dims):
            n *= dict(dims)[d]
        return n * (8 if v.typ == "d" else 4)

    recs = [v for v in variables if is_rec(v)]
    fixed = [v for v in variables if not is_rec(v)]
    order = fixed + recs

    def header(begins):
        h = b"CDF\x02" + struct.pack(">i", nrec)
        h += struct.pack(">ii", 0x0A, len(dims)) if dims else struct.pack(">ii", 0, 0)
        for n, l in dims:
            h += name(n) + struct.pack(">i", 0 if l is None else l)
        h += atts(gatts or {})
        h += struct.pack(">ii", 0x0B, len(order))
        for v in order:
            vs = slab(v)
            vsize = vs if (is_rec(v) and len(recs) == 1) else vs + (-vs % 4)
            h += name(v.name) + struct.pack(">i", len(v.dims))
            h += b"".join(struct.pack(">i", dim_ix[d]) for d in v.dims)
            h += atts(v.attrs) + struct.pack(">ii", _NC_DOUBLE if v.typ == "d" else _NC_INT, vsize)
            h += struct.pack(">q", begins.get(v.name, 0))
        return h

    hsize = len(header({}))
    begins, off = {}, hsize
    for v in fixed:
        begins[v.name] = off
        off += slab(v) + (-slab(v) % 4)
    rec_start = off
    for v in recs:
        begins[v.name] = off
        off += slab(v) if len(recs) == 1 else slab(v) + (-slab(v) % 4)
    be = {"d": ">f8", "i": ">i4"}
    with open(path, "wb") as fh:
        fh.write(header(begins))
        for v in fixed:
            b = np.ascontiguousarray(v.data).astype(be[v.typ]).tobytes()
            fh.write(pad4(b))
        for r in range(nrec):
            for v in recs:
                b = np.ascontiguousarray(v.data[r]).astype(be[v.typ]).tobytes()
                fh.write(b if len(recs) == 1 else pad4(b))
    assert rec_start >= hsize


def _mesh_vars(mesh: Mesh, nC_mem=None):
    """Dimensions and variables of setup_mesh_in_netcdf_file."""
    from .mesh import circumcentres
    nV, nT = mesh.nV, mesh.nT
    cnt = np.diff(mesh.C_ptr)
    nC_mem = int(nC_mem or max(32, int(cnt.max())))
    dims = [("vi", nV), ("ti", nT), ("ci", nC_mem), ("two", 2), ("three", 3)]
    span = max(mesh.xmax - mesh.xmin, mesh.ymax - mesh.ymin)
    scal = {"xmin": mesh.xmin, "xmax": mesh.xmax, "ymin": mesh.ymin, "ymax": mesh.ymax,
            "tol_dist": 1e-9 * span, "lambda_M": 0.0, "phi_M": 0.0, "beta_stereo": 0.0}
    vs = [_Var(k, (), "d", scal[k]) for k in MESH_SCALARS]
    geom = vertex_geometry(mesh)
    src = np.repeat(np.arange(nV), cnt)
    dist = np.hypot(*(mesh.V[mesh.C_ind] - mesh.V[src]).T)
    R = np.full(nV, np.inf)
    np.minimum.at(R, src, dist)
    vs += [
        _Var("V", ("two", "vi"), "d", mesh.V.T, long_name="Vertex coordinates", units="m"),
        _Var("nC", ("vi",), "i", cnt, long_name="Number of vertex-vertex connections"),
        _Var("C", ("ci", "vi"), "i", _ragged_to_padded(mesh.C_ptr, mesh.C_ind, nC_mem).T,
             long_name="Vertex-vertex connections"),
        _Var("niTri", ("vi",), "i", np.diff(mesh.iTri_ptr), long_name="Number of vertex-triangle connections"),
        _Var("iTri", ("ci", "vi"), "i", _ragged_to_padded(mesh.iTri_ptr, mesh.iTri_ind, nC_mem).T,
             long_name="Vertex-triangle connections"),
        _Var("VBI", ("vi",), "i", mesh.VBI, long_name="Vertex boundary index"),
        _Var("Tri", ("three", "ti"), "i", (mesh.Tri + 1).T, long_name="Vertex indices per triangle"),
        _Var("Tricc", ("two", "ti"), "d", circumcentres(mesh).T, long_name="Triangle circumcentre coordinates",
             units="m"),
        _Var("TriC", ("three", "ti"), "i", (np.maximum(mesh.TriC, -1) + 1).T,
             long_name="Triangle-triangle connections"),
        _Var("TriBI", ("ti",), "i", mesh.TriBI, long_name="Triangle boundary index"),
        _Var("TriGC", ("two", "ti"), "d", mesh.TriGC.T, long_name="Triangle geometric centre coordinates",
             units="m"),
        _Var("A", ("vi",), "d", geom.A, long_name="Voronoi cell area", units="m^2"),
        _Var("R", ("vi",), "d", R, long_name="Resolution", units="m"),
    ]
    return dims, vs


def _no_clobber(path):
    if os.path.exists(path):
        raise MeshFileError(f"{path} exists already")


def write_mesh(path, mesh: Mesh, nC_mem=None):
    """A NetCDF mesh file of the reference's layout (classic, 64-bit offsets)."""
    _no_clobber(path)
    dims, vs = _mesh_vars(mesh, nC_mem)
    _write_classic(path, dims, vs)


# --------------------------------------------------------------------------
# velocity restarts
# --------------------------------------------------------------------------

def create_restart(path, mesh: Mesh, stress_balance="SSA", zeta=None):
    """An empty restart file: mesh, time (unlimited), zeta (DIVA) and the
    solver's fields (create_restart_file_SSA / _DIVA)."""
    names = SSA_RESTART_FIELDS if stress_balance == "SSA" else DIVA_RESTART_FIELDS
    if stress_balance != "SSA" and zeta is None:
        raise MeshFileError("a DIVA restart needs the zeta grid")
    _no_clobber(path)
    dims, vs = _mesh_vars(mesh)
    dims = [("time", None)] + dims
    vs.append(_Var("time", ("time",), "d", np.zeros(0), long_name="Time", units="years"))
    if zeta is not None:
        dims.append(("zeta", len(zeta)))
        vs.append(_Var("zeta", ("zeta",), "d", zeta, long_name="Scaled vertical coordinate", units="0-1"))
    nz = len(zeta) if zeta is not None else 0
    for n in names:
        shape = (0, nz, mesh.nT) if n == "eta_3D_b" else (0, mesh.nT)
        d = ("time", "zeta", "ti") if n == "eta_3D_b" else ("time", "ti")
        vs.append(_Var(n, d, "d", np.zeros(shape), units=FIELD_UNITS[n]))
    _write_classic(path, dims, vs)


def write_restart(path, time, fields: dict):
    """Append one time slice (write_to_restart_file_SSA / _DIVA).  3-D fields
    come in the engine's (nz*nT,) triangle-fastest layout.  The file is
    rewritten with the new record and moved over the old one."""
    from scipy.io import netcdf_file
    src = netcdf_file(path, "r", mmap=False)
    try:
        dims = [(k, v) for k, v in src.dimensions.items()]
        rec = [k for k, v in dims if v is None]
        nT = src.dimensions["ti"]
        for n in fields:
            if n not in src.variables:
                raise MeshFileError(f"{path} has no field {n}")
        vs = []
        for k, v in src.variables.items():
            data = np.array(v.data, copy=True)
            attrs = {a: (val.decode() if isinstance(val, bytes) else val) for a, val in v._attributes.items()}
            if v.dimensions and v.dimensions[0] in rec:
                if k == "time":
                    add = np.array([float(time)])
                elif k in fields:
                    a = np.asarray(fields[k], np.float64)
                    add = (a.reshape(-1, nT) if k == "eta_3D_b" else a.reshape(nT))[None]
                else:
                    add = np.zeros((1,) + data.shape[1:], data.dtype)
                data = np.concatenate([data, add.astype(data.dtype)])
            vs.append(_Var(k, v.dimensions, "d" if v.typecode() == "d" else "i", data, **attrs))
    finally:
        src.close()
    tmp = path + ".tmp"
    _write_classic(tmp, dims, vs)
    os.replace(tmp, path)


def read_restart(path, names=None, time=None) -> dict:
    """Fields of one time slice (the last one, or the one at `time`), in the
    engine's layout, with the slice's time under 'time'."""
    f = _open(path)
    try:
        if not f.has("time"):
            raise MeshFileError(f"{path} has no time variable")
        times = np.atleast_1d(f.read("time"))
        if times.size == 0:
            raise MeshFileError(f"{path} holds no time slice")
        it = times.size - 1 if time is None else int(np.argmin(np.abs(times - time)))
        if time is not None and abs(times[it] - time) > 1e-9 * max(1.0, abs(time)):
            raise MeshFileError(f"{path} has no time slice at {time}")
        if names is None:
            names = [n for n in SSA_RESTART_FIELDS + DIVA_RESTART_FIELDS if f.has(n)]
        out = {"time": float(times[it])}
        for n in names:
            a = f.read(n)[it]
            out[n] = np.ascontiguousarray(a.reshape(-1), np.float64)
        return out
    finally:
        f.close()
