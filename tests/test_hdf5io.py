"""HDF5 files either side of the time loop (host only, no GPU): receiver
files in the ReceiverHdf5::write layout (ReceiverHdf5.cpp:31-87), source
files read the way SourceHdf5::loadData reads them (SourceHdf5.cpp:40-50)
and movie frames appended the way Problem::saveSolution is called
(Problem.cpp:217-237, after every step).
The files are checked with libhdf5 called directly from the test, not
through the library under test."""
import ctypes as C
import os

import numpy as np
import pytest

from salvus_amd import sem

HID = C.c_int64
DEFAULT = HID(0)


def _libhdf5():
    for name in (os.environ.get("SEM_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103",
                 "/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103"):
        if not name:
            continue
        try:
            h = C.CDLL(name)
        except OSError:
            continue
        h.H5open()
        h.H5Fopen.restype = HID
        h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, HID]
        h.H5Fcreate.restype = HID
        h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, HID, HID]
        h.H5Fclose.argtypes = [HID]
        h.H5Gcreate2.restype = HID
        h.H5Gcreate2.argtypes = [HID, C.c_char_p, HID, HID, HID]
        h.H5Gclose.argtypes = [HID]
        h.H5Dopen2.restype = HID
        h.H5Dopen2.argtypes = [HID, C.c_char_p, HID]
        h.H5Dcreate2.restype = HID
        h.H5Dcreate2.argtypes = [HID, C.c_char_p, HID, HID, HID, HID, HID]
        h.H5Dread.argtypes = [HID, HID, HID, HID, HID, C.c_void_p]
        h.H5Dwrite.argtypes = [HID, HID, HID, HID, HID, C.c_void_p]
        h.H5Dclose.argtypes = [HID]
        h.H5Dget_space.restype = HID
        h.H5Dget_space.argtypes = [HID]
        h.H5Dget_type.restype = HID
        h.H5Dget_type.argtypes = [HID]
        h.H5Tget_size.restype = C.c_size_t
        h.H5Tget_size.argtypes = [HID]
        h.H5Tget_class.argtypes = [HID]
        h.H5Tclose.argtypes = [HID]
        h.H5Sget_simple_extent_ndims.argtypes = [HID]
        h.H5Sget_simple_extent_dims.argtypes = [HID, C.POINTER(C.c_ulonglong), C.c_void_p]
        h.H5Screate_simple.restype = HID
        h.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]
        h.H5Sclose.argtypes = [HID]
        h.H5Lexists.argtypes = [HID, C.c_char_p, HID]
        h.H5Eset_auto2.argtypes = [HID, C.c_void_p, C.c_void_p]
        h.H5Eset_auto2(DEFAULT, None, None)
        h.float_t = HID.in_dll(h, "H5T_NATIVE_FLOAT_g")
        h.double_t = HID.in_dll(h, "H5T_NATIVE_DOUBLE_g")
        h.int_t = HID.in_dll(h, "H5T_NATIVE_INT_g")
        h.H5Aopen.restype = HID
        h.H5Aopen.argtypes = [HID, C.c_char_p, HID]
        h.H5Aread.argtypes = [HID, HID, C.c_void_p]
        h.H5Aclose.argtypes = [HID]
        return h
    return None


H5 = _libhdf5()
pytestmark = pytest.mark.skipif(H5 is None, reason="libhdf5 not found on this host")


def read_dataset(path, name, dtype=np.float32):
    f = H5.H5Fopen(str(path).encode(), 0, DEFAULT)
    assert f >= 0
    d = H5.H5Dopen2(f, name.encode(), DEFAULT)
    assert d >= 0, name
    sp = H5.H5Dget_space(d)
    nd = H5.H5Sget_simple_extent_ndims(sp)
    dims = (C.c_ulonglong * max(nd, 1))()
    H5.H5Sget_simple_extent_dims(sp, dims, None)
    ty = H5.H5Dget_type(d)
    size, cls = H5.H5Tget_size(ty), H5.H5Tget_class(ty)
    out = np.zeros([dims[i] for i in range(nd)], dtype=dtype)
    mem_t = {np.float64: H5.double_t, np.int32: H5.int_t}.get(dtype, H5.float_t)
    assert H5.H5Dread(d, mem_t, DEFAULT, DEFAULT, DEFAULT, out.ctypes.data) >= 0
    H5.H5Tclose(ty)
    H5.H5Sclose(sp)
    H5.H5Dclose(d)
    H5.H5Fclose(f)
    return out, size, cls


def exists(path, name):
    f = H5.H5Fopen(str(path).encode(), 0, DEFAULT)
    r = H5.H5Lexists(f, name.encode(), DEFAULT)
    H5.H5Fclose(f)
    return r > 0


def write_source_file(path, datasets):
    """datasets: {source name: 1-D float64 array} written as <name>/data."""
    f = H5.H5Fcreate(str(path).encode(), 2, DEFAULT, DEFAULT)
    for name, arr in datasets.items():
        g = H5.H5Gcreate2(f, name.encode(), DEFAULT, DEFAULT, DEFAULT)
        H5.H5Gclose(g)
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        dims = (C.c_ulonglong * 1)(arr.size)
        sp = H5.H5Screate_simple(1, dims, None)
        d = H5.H5Dcreate2(f, ("%s/data" % name).encode(), H5.double_t, sp, DEFAULT, DEFAULT, DEFAULT)
        assert H5.H5Dwrite(d, H5.double_t, DEFAULT, DEFAULT, DEFAULT, arr.ctypes.data) >= 0
        H5.H5Dclose(d)
        H5.H5Sclose(sp)
    H5.H5Fclose(f)


def test_receiver_file_layout(tmp_path):
    rng = np.random.default_rng(3)
    seis = rng.standard_normal((37, 3, 2)).astype(np.float32)  # nsamples x nrec x c
    path = tmp_path / "receivers.h5"
    sem.write_receivers_h5(path, ["r0", "r1", "sta_2"], seis)
    for r, name in enumerate(["r0", "r1", "sta_2"]):
        assert exists(path, name)
        for k, field in enumerate(("ux", "uy")):
            arr, size, cls = read_dataset(path, "/%s/%s" % (name, field))
            assert arr.shape == (37,) and size == 4 and cls == 1  # H5T_FLOAT, 4 bytes
            np.testing.assert_array_equal(arr, seis[:, r, k])
    assert not exists(path, "r0/uz")


def test_receiver_file_append_and_errors(tmp_path):
    path = tmp_path / "receivers.h5"
    a = np.arange(10, dtype=np.float32).reshape(10, 1, 1)
    sem.write_receivers_h5(path, ["a"], a)
    sem.write_receivers_h5(path, ["b"], 2 * a, append=True)
    np.testing.assert_array_equal(read_dataset(path, "/a/u")[0], a[:, 0, 0])
    np.testing.assert_array_equal(read_dataset(path, "/b/u")[0], 2 * a[:, 0, 0])
    with pytest.raises(sem.SemError, match="already in the file"):
        sem.write_receivers_h5(path, ["a"], a, append=True)
    with pytest.raises(sem.SemError, match="no '/'"):
        sem.write_receivers_h5(path, ["x/y"], a)
    sem.write_receivers_h5(path, ["c"], a)  # no append: a fresh file
    assert exists(path, "c") and not exists(path, "a")
    three = np.ones((4, 2, 3), dtype=np.float32)
    sem.write_receivers_h5(tmp_path / "e3.h5", ["p", "q"], three)
    assert exists(tmp_path / "e3.h5", "q/uz")


def test_source_file_column_major(tmp_path):
    nsteps, ncomp = 25, 3
    mat = np.arange(nsteps * ncomp, dtype=np.float64).reshape(nsteps, ncomp) * 0.5 - 3.0
    # stored the way the reference's Eigen (column-major) matrix is filled
    path = tmp_path / "sources.h5"
    write_source_file(path, {"src0": mat.T.ravel(), "short": np.zeros(7)})
    got = sem.read_source_h5(path, "src0", nsteps, ncomp)
    np.testing.assert_array_equal(got, mat)  # row n = fire(t, n)
    with pytest.raises(sem.SemError, match="expected nsteps x ncomp"):
        sem.read_source_h5(path, "short", nsteps, ncomp)
    with pytest.raises(sem.SemError, match="Can't read data for source 'missing'"):
        sem.read_source_h5(path, "missing", nsteps, ncomp)
    with pytest.raises(sem.SemError, match="cannot open"):
        sem.read_source_h5(tmp_path / "nope.h5", "src0", nsteps, ncomp)


def test_movie_frames(tmp_path):
    rng = np.random.default_rng(5)
    path = tmp_path / "movie.h5"
    frames = [{"ux": rng.standard_normal(11), "uy": rng.standard_normal(11)} for _ in range(3)]
    times = [0.1, 0.2, 0.30000000000000004]
    for i, (fr, t) in enumerate(zip(frames, times)):
        sem.write_movie_frame(path, fr, t, append=i > 0)
    for name in ("ux", "uy"):
        arr, size, cls = read_dataset(path, "/fields/" + name, np.float64)
        assert arr.shape == (3, 11) and size == 8 and cls == 1  # H5T_FLOAT, 8 bytes
        np.testing.assert_array_equal(arr, np.stack([fr[name] for fr in frames]))
    t, size, _ = read_dataset(path, "/time", np.float64)
    assert size == 8
    np.testing.assert_array_equal(t, times)
    # a frame must match what the file holds: same length, every field present
    with pytest.raises(sem.SemError, match="do not match this frame"):
        sem.write_movie_frame(path, {"ux": np.zeros(12)}, 0.4)
    with pytest.raises(sem.SemError, match="do not match this frame"):
        sem.write_movie_frame(path, {"vx": np.zeros(11)}, 0.4)
    assert read_dataset(path, "/time", np.float64)[0].shape == (3,)
    with pytest.raises(sem.SemError, match="no '/'"):
        sem.write_movie_frame(path, {"a/b": np.zeros(11)}, 0.4)
    sem.write_movie_frame(path, {"u": np.arange(4.0)}, 1.5, append=False)  # a fresh file
    assert not exists(path, "fields/ux")
    np.testing.assert_array_equal(read_dataset(path, "/fields/u", np.float64)[0], [np.arange(4.0)])


def test_movie_field_names():
    assert set(sem.movie_field_names(1)) == {"u", "v", "a", "a_"}
    names = sem.movie_field_names(3)
    assert set(names) == {p + s + q for p in "uva" for s in "xyz" for q in ("",)} | {"ax_", "ay_", "az_"}
    assert names["vy"] == (sem.FIELD["v"], 1) and names["az_"] == names["az"] == (sem.FIELD["a_"], 2)


def test_movie_mesh_then_frames(tmp_path):
    """The mesh goes into the movie file before any frame, as the reference
    DMViews it into the viewer (Problem.cpp:133-137): vertices, cells
    (cell_dim attribute) under PETSc's DMPlex view names, and the DOF
    coordinates that place the frames' values; frames then append."""
    from oracle import pyoracle as po
    from salvus_amd import meshgen
    cells, coords = meshgen.box_quad(3, 2, 30.0, 20.0)
    S = po.Sim(2, 4, 0, cells, coords, np.full(cells.size, 3000.0))
    xyz = S.node_coords()
    path = tmp_path / "movie.h5"
    sem.write_movie_mesh(path, 2, cells, coords, xyz)
    for k in range(2):
        sem.write_movie_frame(path, {"u": np.full(S.G, float(k))}, 0.1 * (k + 1), append=True)
    v, size, cls = read_dataset(path, "/geometry/vertices", np.float64)
    np.testing.assert_array_equal(v, coords)
    c, size, cls = read_dataset(path, "/topology/cells", np.int32)
    assert size == 4 and cls == 0  # H5T_INTEGER, 4 bytes
    np.testing.assert_array_equal(c, cells)
    n, _, _ = read_dataset(path, "/geometry/nodes", np.float64)
    np.testing.assert_array_equal(n, xyz)
    u, _, _ = read_dataset(path, "/fields/u", np.float64)
    assert u.shape == (2, S.G) and u[1, 0] == 1.0
    f = H5.H5Fopen(str(path).encode(), 0, DEFAULT)
    d = H5.H5Dopen2(f, b"/topology/cells", DEFAULT)
    a = H5.H5Aopen(d, b"cell_dim", DEFAULT)
    cd = np.zeros(1, dtype=np.int32)
    assert a >= 0 and H5.H5Aread(a, H5.int_t, cd.ctypes.data) >= 0 and cd[0] == 2
    H5.H5Aclose(a)
    H5.H5Dclose(d)
    H5.H5Fclose(f)
    with pytest.raises(sem.SemError, match="out of range"):
        sem.write_movie_mesh(path, 2, cells + 100, coords, xyz)
