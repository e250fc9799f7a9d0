"""Mesh and restart files in the reference's NetCDF layout
(netcdf_setup_grid_mesh_in_file.f90:130-453, netcdf_setup_grid_mesh_from_file.f90:133-241,
SSA_main.f90:538-627, DIVA_main.f90:808-910): the dimension names and their
Fortran (fastest-first) order, 1-based indices with 0 padding, round trips
through classic and NetCDF-4 (HDF5) files, the reader's consistency checks,
and operators built from a mesh read back."""

import ctypes as C
import os

import numpy as np
import pytest

import __graft_entry__ as ge

ge.load_package()
from ufemism2_amd import netcdf_io as N, workloads as W  # noqa: E402
from ufemism2_amd.mesh import uniform_mesh  # noqa: E402
from ufemism2_amd.operators import build_operators  # noqa: E402


@pytest.fixture(scope="module")
def mesh():
    return W.channel_small(target_nV=1500).mesh


def _same_mesh(a, b):
    assert (a.xmin, a.xmax, a.ymin, a.ymax) == (b.xmin, b.xmax, b.ymin, b.ymax)
    for k in ("V", "Tri", "TriC", "VBI", "TriBI", "TriGC", "C_ptr", "C_ind", "iTri_ptr", "iTri_ind"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_mesh_file_layout(tmp_path, mesh):
    from scipy.io import netcdf_file
    p = str(tmp_path / "mesh.nc")
    N.write_mesh(p, mesh)
    with open(p, "rb") as fh:
        assert fh.read(4) == b"CDF\x02"
    f = netcdf_file(p, "r", mmap=False)
    try:
        d = f.dimensions
        assert (d["vi"], d["ti"], d["two"], d["three"]) == (mesh.nV, mesh.nT, 2, 3) and d["ci"] >= 32
        assert f.variables["V"].dimensions == ("two", "vi")
        assert f.variables["C"].dimensions == ("ci", "vi")
        assert f.variables["Tri"].dimensions == ("three", "ti")
        Tri = f.variables["Tri"].data
        assert Tri.min() == 1 and Tri.max() == mesh.nV
        Cv = np.array(f.variables["C"].data)
        nC = np.array(f.variables["nC"].data)
        assert (Cv[nC[0]:, 0] == 0).all() and (Cv[:nC[0], 0] > 0).all()
        TriC = np.array(f.variables["TriC"].data)
        assert (TriC == 0).sum() == (mesh.TriC < 0).sum()
        for k in N.MESH_SCALARS:
            assert f.variables[k].shape == ()
        assert abs(np.array(f.variables["A"].data).sum() - (mesh.xmax - mesh.xmin) * (mesh.ymax - mesh.ymin)) < 1e-3
    finally:
        f.close()


def test_mesh_round_trip_classic(tmp_path, mesh):
    p = str(tmp_path / "mesh.nc")
    N.write_mesh(p, mesh)
    m2 = N.read_mesh(p)
    _same_mesh(mesh, m2)
    assert m2.digest() == mesh.digest()
    with pytest.raises(N.MeshFileError):
        N.write_mesh(p, mesh)  # no clobbering, as the reference's NF90_NOCLOBBER


def test_operators_from_a_mesh_read_back(tmp_path):
    m = uniform_mesh(0.0, 100e3, 0.0, 60e3, 8e3, seed=4)
    p = str(tmp_path / "m.nc")
    N.write_mesh(p, m)
    a, b = build_operators(m), build_operators(N.read_mesh(p))
    for fam in ("ab", "ba", "bb"):
        fa, fb = getattr(a, fam), getattr(b, fam)
        assert np.array_equal(fa.ptr, fb.ptr) and np.array_equal(fa.ind, fb.ind), fam
        for k, v in fa.vals.items():
            assert np.array_equal(v, fb.vals[k]), (fam, k)


def test_inconsistent_mesh_is_rejected(tmp_path, mesh):
    from scipy.io import netcdf_file
    p = str(tmp_path / "bad.nc")
    N.write_mesh(p, mesh)
    f = netcdf_file(p, "a", mmap=False)
    c = f.variables["C"]
    col = np.array(c[:, 5])
    col[0], col[1] = col[1], col[0] + 0  # still a permutation: allowed
    c[:, 5] = col
    f.close()
    N.read_mesh(p)
    f = netcdf_file(p, "a", mmap=False)
    c = f.variables["C"]
    col = np.array(c[:, 5])
    col[0] = (col[0] % mesh.nV) + 1 if col[0] != col[1] + 1 else col[0]
    c[:, 5] = col
    f.close()
    with pytest.raises(N.MeshFileError):
        N.read_mesh(p)


def test_ssa_restart_round_trip(tmp_path, mesh):
    p = str(tmp_path / "restart_ice_velocity_SSA_00001.nc")
    N.create_restart(p, mesh, "SSA")
    rng = np.random.default_rng(1)
    u1, v1 = rng.normal(size=mesh.nT), rng.normal(size=mesh.nT)
    u2, v2 = rng.normal(size=mesh.nT), rng.normal(size=mesh.nT)
    N.write_restart(p, 10.0, {"u_b": u1, "v_b": v1})
    N.write_restart(p, 20.0, {"u_b": u2, "v_b": v2})
    last = N.read_restart(p)
    assert last["time"] == 20.0 and np.array_equal(last["u_b"], u2) and np.array_equal(last["v_b"], v2)
    first = N.read_restart(p, time=10.0)
    assert np.array_equal(first["u_b"], u1)
    with pytest.raises(N.MeshFileError):
        N.read_restart(p, time=15.0)
    _same_mesh(mesh, N.read_mesh(p))


def test_diva_restart_round_trip(tmp_path, mesh):
    p = str(tmp_path / "restart_ice_velocity_DIVA_00001.nc")
    zeta = np.linspace(0.0, 1.0, 10)
    N.create_restart(p, mesh, "DIVA", zeta=zeta)
    rng = np.random.default_rng(2)
    st = {n: rng.normal(size=mesh.nT * (10 if n == "eta_3D_b" else 1)) for n in N.DIVA_RESTART_FIELDS}
    N.write_restart(p, 0.0, st)
    got = N.read_restart(p)
    for n in N.DIVA_RESTART_FIELDS:
        assert np.array_equal(got[n], st[n]), n
    from scipy.io import netcdf_file
    f = netcdf_file(p, "r", mmap=False)
    assert f.variables["eta_3D_b"].dimensions == ("time", "zeta", "ti")
    assert np.array_equal(np.array(f.variables["eta_3D_b"][0]), st["eta_3D_b"].reshape(10, mesh.nT))
    f.close()


def _write_hdf5_like_netcdf4(path, arrays):
    """Datasets named as the variables, C-order shapes as the netCDF library
    stores them (what a NetCDF-4 file holds, minus its dimension scales)."""
    x = N._Hdf5()
    h = x.h
    fid = h.H5Fcreate(path.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    assert fid >= 0
    for name, a in arrays.items():
        a = np.ascontiguousarray(a)
        is_int = a.dtype.kind in "iu"
        a = a.astype(np.int32 if is_int else np.float64)
        if a.ndim == 0:
            sp = h.H5Screate(0)  # H5S_SCALAR
        else:
            dims = (C.c_uint64 * a.ndim)(*a.shape)
            sp = h.H5Screate_simple(a.ndim, dims, None)
        t = x.int32 if is_int else x.double
        d = h.H5Dcreate2(fid, name.encode(), t, sp, 0, 0, 0)
        assert d >= 0
        assert h.H5Dwrite(d, t, 0, 0, 0, a.ctypes.data_as(C.c_void_p)) >= 0
        h.H5Dclose(d)
        h.H5Sclose(sp)
    h.H5Fclose(fid)


def test_netcdf4_hdf5_mesh_is_read(tmp_path, mesh):
    try:
        N._Hdf5()
    except N.MeshFileError:
        pytest.skip("no HDF5 library on this machine")
    from scipy.io import netcdf_file
    p = str(tmp_path / "mesh.nc")
    N.write_mesh(p, mesh)
    f = netcdf_file(p, "r", mmap=False)
    arrays = {k: np.array(v.data).reshape(v.shape) for k, v in f.variables.items()}
    f.close()
    q = str(tmp_path / "mesh_nc4.nc")
    _write_hdf5_like_netcdf4(q, arrays)
    with open(q, "rb") as fh:
        assert fh.read(8) == b"\x89HDF\r\n\x1a\n"
    _same_mesh(mesh, N.read_mesh(q))


def test_not_a_netcdf_file(tmp_path):
    p = str(tmp_path / "x.nc")
    with open(p, "wb") as fh:
        fh.write(b"hello world")
    with pytest.raises(N.MeshFileError):
        N.read_mesh(p)
    assert not os.path.exists(str(tmp_path / "y.nc"))
