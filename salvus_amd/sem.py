"""Python front end of the MI355X SEM library (salvus_amd/_lib/libsalvus_sem.so).

Thin ctypes bindings over the C ABI declared in include/salvus_sem.h, with
classes named after the reference's plugin/operator API:

  Mesh            Mesh::read(dim, ...) + setupGlobalDof + element set-up
                  (Mesh.cpp:31-51, :361-452; Order2Newmark.cpp:32-66)
  Batch           shots advanced together through the fused or plain path
  Order2Newmark   one shot driven like the reference's Problem loop
                  (Problem.cpp:143-215, Order2Newmark.cpp:68-115, Main.cpp:45-66)

The HIP library must be built (python -c "import __graft_entry__ as g;
g.build()"); nothing here falls back to a CPU path: a missing library or a
missing GPU raises.
"""
import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libsalvus_sem.so"

ACOUSTIC, ELASTIC = 0, 1
QUIRK_NEWMARK3D, QUIRK_E3D_SOURCE, QUIRK_QUAD_JACOBIAN = 1, 2, 4
DEFAULT_QUIRKS = QUIRK_E3D_SOURCE | QUIRK_QUAD_JACOBIAN
PATH_PLAIN, PATH_FUSED = 0, 1
FIELD = {"u": 0, "v": 1, "a_": 2}

# Every entry point of include/salvus_sem.h.
EXPORTS = (
    "sem_last_error", "sem_device_count", "sem_topology_build", "sem_mesh_create",
    "sem_mesh_destroy", "sem_mesh_info", "sem_mesh_get_conn", "sem_mesh_get_colours",
    "sem_mesh_get_inverse_mass", "sem_mesh_get_node_coords", "sem_mesh_set_dirichlet",
    "sem_mesh_apply_stiffness", "sem_mesh_check_colouring", "sem_batch_create",
    "sem_batch_destroy", "sem_batch_add_shot", "sem_shot_add_source",
    "sem_shot_add_moment_source", "sem_shot_add_receiver", "sem_shot_set_field",
    "sem_shot_get_field", "sem_batch_prepare", "sem_batch_run", "sem_shot_progress",
    "sem_shot_get_seismograms", "sem_batch_config", "sem_batch_lanes", "sem_batch_kernel_time",
    "sem_reference_step_count", "sem_server_create", "sem_server_destroy", "sem_server_clock",
    "sem_server_submit", "sem_server_pump", "sem_server_request", "sem_server_result",
    "sem_mesh_autotune", "sem_shot_add_file_source", "sem_receivers_write_h5",
    "sem_source_read_h5", "sem_movie_write_h5", "sem_mesh_apply_stiffness_noalloc",
    "sem_server_cancel", "sem_movie_write_mesh_h5", "sem_coupling_create",
    "sem_coupling_destroy", "sem_coupling_info", "sem_batch_couple", "sem_server_set_grouping",
    "sem_server_request_group",
)
SOURCE_DOUBLES = 19

_lib = None
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_VP = C.c_void_p


class SemError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SemError("HIP library %s is not built; run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(str(LIB_PATH))
        L.sem_last_error.restype = C.c_char_p
        L.sem_device_count.argtypes = [C.POINTER(C.c_int)]
        L.sem_topology_build.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, _I32, _I32, _I32,
                                         _I64]
        L.sem_mesh_create.argtypes = [C.POINTER(_VP), C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int64, C.c_int64, _I32, _D, _D, C.c_int]
        L.sem_mesh_destroy.argtypes = [_VP]
        L.sem_mesh_info.argtypes = [_VP, _I64]
        L.sem_mesh_get_conn.argtypes = [_VP, _I32]
        L.sem_mesh_get_colours.argtypes = [_VP, _I32]
        L.sem_mesh_get_inverse_mass.argtypes = [_VP, _F]
        L.sem_mesh_get_node_coords.argtypes = [_VP, _D]
        L.sem_mesh_set_dirichlet.argtypes = [_VP, C.POINTER(C.c_uint8)]
        L.sem_mesh_apply_stiffness.argtypes = [_VP, C.c_int, _F, _F]
        L.sem_mesh_check_colouring.argtypes = [_VP, _I64]
        L.sem_mesh_apply_stiffness_noalloc.argtypes = [_VP, C.c_void_p, C.c_void_p, C.c_void_p]
        L.sem_batch_create.argtypes = [C.POINTER(_VP), C.c_int]
        L.sem_batch_destroy.argtypes = [_VP]
        L.sem_batch_add_shot.argtypes = [_VP, _VP, C.c_int64, C.c_double, C.POINTER(C.c_int)]
        L.sem_shot_add_source.argtypes = [_VP, C.c_int, _D, C.c_double, C.c_double, C.c_double,
                                          _D]
        L.sem_shot_add_moment_source.argtypes = [_VP, C.c_int, _D, _D, C.c_double, C.c_double,
                                                 C.c_double]
        L.sem_shot_add_receiver.argtypes = [_VP, C.c_int, _D]
        L.sem_shot_set_field.argtypes = [_VP, C.c_int, C.c_int, _D]
        L.sem_shot_get_field.argtypes = [_VP, C.c_int, C.c_int, _D]
        L.sem_batch_prepare.argtypes = [_VP, C.c_int, C.c_int, C.c_int]
        L.sem_batch_run.argtypes = [_VP, C.c_int64]
        L.sem_shot_progress.argtypes = [_VP, C.c_int, _I64, _D]
        L.sem_shot_get_seismograms.argtypes = [_VP, C.c_int, _F]
        L.sem_batch_config.argtypes = [_VP, _I64]
        L.sem_batch_lanes.argtypes = [_VP, C.POINTER(C.c_int)]
        L.sem_batch_kernel_time.argtypes = [_VP, _D, _I64]
        L.sem_reference_step_count.restype = C.c_int64
        L.sem_reference_step_count.argtypes = [C.c_double, C.c_double, _D]
        L.sem_server_create.argtypes = [C.POINTER(_VP), C.c_int, C.c_int, C.c_int64, C.c_int,
                                        C.c_int]
        L.sem_server_destroy.argtypes = [_VP]
        L.sem_server_clock.argtypes = [_VP, _D]
        L.sem_server_submit.argtypes = [_VP, _VP, C.c_int64, C.c_double, C.c_double, C.c_int, _D,
                                        C.c_int, _D, C.POINTER(C.c_int)]
        L.sem_server_pump.argtypes = [_VP, _D]
        L.sem_server_request.argtypes = [_VP, C.c_int, _D]
        L.sem_server_result.argtypes = [_VP, C.c_int, _F]
        L.sem_server_cancel.argtypes = [_VP, C.c_int]
        L.sem_server_set_grouping.argtypes = [_VP, C.c_int, C.c_double]
        L.sem_server_request_group.argtypes = [_VP, C.c_int, C.POINTER(C.c_int)]
        L.sem_movie_write_mesh_h5.argtypes = [C.c_char_p, C.c_int, C.c_int64, _D, C.c_int64, _I32,
                                              C.c_int64, _D]
        L.sem_coupling_create.argtypes = [C.POINTER(_VP), _VP, _VP, _D]
        L.sem_coupling_destroy.argtypes = [_VP]
        L.sem_coupling_info.argtypes = [_VP, _I64]
        L.sem_batch_couple.argtypes = [_VP, _VP, C.c_int, C.c_int]
        L.sem_mesh_autotune.argtypes = [_VP, C.c_int, C.c_int, _I64]
        L.sem_shot_add_file_source.argtypes = [_VP, C.c_int, _D, C.c_int, C.c_int64, _D]
        L.sem_receivers_write_h5.argtypes = [C.c_char_p, C.c_int, C.c_int,
                                             C.POINTER(C.c_char_p), C.c_int,
                                             C.POINTER(C.c_char_p), C.c_int64, _F]
        L.sem_source_read_h5.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_int, _D]
        L.sem_movie_write_h5.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_char_p),
                                         C.c_int64, _D, C.c_double]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SemError(lib().sem_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_D)


def _f(a):
    return a.ctypes.data_as(_F)


def _i(a):
    return a.ctypes.data_as(_I32)


def _l(a):
    return a.ctypes.data_as(_I64)


def device_count():
    n = C.c_int(0)
    lib().sem_device_count(C.byref(n))
    return n.value


def reference_step_count(duration, dt):
    """Iterations of the reference's while (t < duration) loop and its dt
    (Options.cpp:84-97, Main.cpp:45)."""
    out = np.zeros(1)
    n = lib().sem_reference_step_count(duration, dt, _d(out))
    return int(n), float(out[0])


FIELD_NAMES = {1: ("u",), 2: ("ux", "uy"), 3: ("ux", "uy", "uz")}


def _cstrings(items):
    arr = (C.c_char_p * len(items))()
    arr[:] = [s.encode() for s in items]
    return arr


def write_receivers_h5(path, names, seis, fields=None, append=False):
    """ReceiverHdf5::write format: /<name>/<field> FP32 datasets.  seis is
    nsamples x nrec x c (Batch.seismograms); fields default to the physics'
    field names (u; ux, uy; ux, uy, uz)."""
    seis = np.ascontiguousarray(seis, dtype=np.float32)
    if seis.ndim != 3 or seis.shape[1] != len(names):
        raise ValueError("seis must be nsamples x len(names) x c")
    fields = tuple(fields) if fields is not None else FIELD_NAMES[seis.shape[2]]
    _check(lib().sem_receivers_write_h5(str(path).encode(), int(bool(append)), len(names),
                                        _cstrings(list(names)), len(fields), _cstrings(list(fields)),
                                        seis.shape[0], _f(seis)))


def read_source_h5(path, name, nsteps, ncomp):
    """SourceHdf5::loadData: dataset <name>/data as an nsteps x ncomp matrix
    (stored column-major, as the reference reads it)."""
    out = np.zeros(int(nsteps) * int(ncomp))
    _check(lib().sem_source_read_h5(str(path).encode(), name.encode(), nsteps, ncomp, _d(out)))
    return out.reshape(ncomp, nsteps).T.copy()


def movie_field_names(c):
    """The reference's global field names of a physics with c components
    (Order2Newmark fields u, v, a, a_; Elastic*D with x, y[, z] suffixes), in
    the order (field, component) they map to: a and a_ are the same values
    after a step (takeTimeStep sets a_ = a)."""
    sfx = ("",) if c == 1 else ("x", "y", "z")[:c]
    return {b + s + t: (FIELD["a_" if b == "a" else b], k)
            for b, t in (("u", ""), ("v", ""), ("a", ""), ("a", "_")) for k, s in enumerate(sfx)}


def write_movie_frame(path, fields, time, append=True):
    """Problem::saveSolution layout: appends one frame, fields = {name:
    values over the global DOFs}, to /fields/<name> [frames][n] (FP64) and
    the time to /time."""
    names = list(fields)
    vals = np.ascontiguousarray(np.stack([np.asarray(fields[k], dtype=np.float64).ravel() for k in names]))
    _check(lib().sem_movie_write_h5(str(path).encode(), int(bool(append)), len(names), _cstrings(names),
                                    vals.shape[1], _d(vals), float(time)))


def write_movie_mesh(path, dim, cells, coords, node_xyz):
    """The movie file's mesh part (written before the frames, as the
    reference DMViews the mesh first, Problem.cpp:133-137): vertices, cells
    and the coordinates of the global DOFs the frames hold."""
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    xyz = np.ascontiguousarray(node_xyz, dtype=np.float64)
    nv = 4 if dim == 2 else 8
    _check(lib().sem_movie_write_mesh_h5(str(path).encode(), dim, coords.size // dim, _d(coords),
                                         cells.size // nv, _i(cells), xyz.size // dim, _d(xyz)))


def topology(dim, order, cells, nvert):
    """Host-only gather map and colouring: (conn[E, nd], colour[E], info)."""
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    nv = 4 if dim == 2 else 8
    E = cells.size // nv
    nd = (order + 1) ** dim
    conn = np.zeros(E * nd, dtype=np.int32)
    col = np.zeros(E, dtype=np.int32)
    info = np.zeros(4, dtype=np.int64)
    _check(lib().sem_topology_build(dim, order, E, nvert, _i(cells), _i(conn), _i(col), _l(info)))
    return conn.reshape(E, nd), col, dict(G=int(info[0]), ncolours=int(info[1]),
                                          faces=int(info[2]), edges=int(info[3]))


class Mesh:
    """Device-resident mesh: gather map, colouring, FP64-built geometry and
    inverse mass (Mesh.cpp:31-51, :361-452; Order2Newmark.cpp:32-66)."""

    def __init__(self, dim, order, physics, cells, coords, vertex_par, quirks=DEFAULT_QUIRKS,
                 device=0):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        par = np.ascontiguousarray(vertex_par, dtype=np.float64)
        nv = 4 if dim == 2 else 8
        self.dim, self.order, self.physics, self.quirks = dim, order, physics, quirks
        self.cells, self.coords = cells, coords  # kept for the movie file's mesh
        self.E = cells.size // nv
        self.V = coords.size // dim
        self._h = _VP()
        _check(lib().sem_mesh_create(C.byref(self._h), device, dim, order, physics, self.E, self.V,
                                     _i(cells), _d(coords), _d(par), quirks))
        self._read_info()

    def _read_info(self):
        info = np.zeros(16, dtype=np.int64)
        lib().sem_mesh_info(self._h, _l(info))
        self.G, self.ncolours, self.c, self.nd = (int(x) for x in info[1:5])
        self.affine = bool(info[5])
        self.device_bytes = int(info[6])
        self.ntiles = int(info[8])
        t = int(info[9])
        self.tile_shape = (t // 10000, (t // 100) % 100, t % 100) if t else None
        self.tile_slots, self.tile_valid_slots = int(info[10]), int(info[11])
        self.tile_acc_reads, self.tile_acc_writes = int(info[12]), int(info[13])
        self.tile_elems = int(info[14])
        self.tile_diagonal = bool(info[15])

    def autotune(self, reps=5, shots=1):
        """Times every compiled tile shape over `reps` steps of one group of
        `shots` shots (the batch size the mesh will serve) and keeps the
        fastest (call before the mesh joins a batch); returns the shape."""
        v = np.zeros(1, dtype=np.int64)
        _check(lib().sem_mesh_autotune(self._h, reps, shots, _l(v)))
        self._read_info()
        return self.tile_shape

    def close(self):
        if getattr(self, "_h", None):
            lib().sem_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def conn(self):
        out = np.zeros(self.E * self.nd, dtype=np.int32)
        _check(lib().sem_mesh_get_conn(self._h, _i(out)))
        return out.reshape(self.E, self.nd)

    def colours(self):
        out = np.zeros(self.E, dtype=np.int32)
        _check(lib().sem_mesh_get_colours(self._h, _i(out)))
        return out

    def inverse_mass(self):
        out = np.zeros(self.G, dtype=np.float32)
        _check(lib().sem_mesh_get_inverse_mass(self._h, _f(out)))
        return out

    def node_coords(self):
        out = np.zeros(self.G * self.dim)
        _check(lib().sem_mesh_get_node_coords(self._h, _d(out)))
        return out.reshape(self.G, self.dim)

    def set_dirichlet(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(lib().sem_mesh_set_dirichlet(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def apply_stiffness(self, u, path=PATH_FUSED):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        out = np.zeros(self.c * self.G, dtype=np.float32)
        _check(lib().sem_mesh_apply_stiffness(self._h, path, _f(u), _f(out)))
        return out.reshape(self.c, self.G)

    def check_colouring(self):
        v = np.zeros(1, dtype=np.int64)
        _check(lib().sem_mesh_check_colouring(self._h, _l(v)))
        return int(v[0])


def split_coupled(cells, coords, fluid, vertex_par):
    """The fluid and the solid part of a mixed quad mesh (the reference's
    single mesh whose elements are Scalar or Elastic2D, Mesh.cpp:480-521), as
    the inputs of two meshes: returns ((cells, coords, vertex_par) of the
    fluid cells, the same of the solid cells), each with its vertices
    renumbered in their original order.  `vertex_par` is a pair (fluid
    cells x 4 x 1 VP, solid cells x 4 x 5) or None."""
    cells = np.asarray(cells)
    coords = np.asarray(coords, dtype=np.float64)
    fluid = np.asarray(fluid, dtype=bool)
    if cells.ndim != 2 or cells.shape[1] != 4 or fluid.shape != (len(cells),):
        raise SemError("split_coupled needs quad cells (E x 4) and one flag per cell")
    out = []
    for k, sel in enumerate((fluid, ~fluid)):
        sub = cells[sel]
        used = np.unique(sub)
        remap = np.full(len(coords), -1, dtype=np.int64)
        remap[used] = np.arange(len(used))
        par = None if vertex_par is None else vertex_par[k]
        out.append((remap[sub].astype(np.int32), coords[used], par))
    return out[0], out[1]


class Coupling:
    """2-D fluid-solid coupling of an acoustic and an elastic mesh cut from
    one quad mesh (the coupled elements of Mesh.cpp:480-521 with the edge
    integrals of ElasticToAcoustic.cpp:17-41 and AcousticToElastic2D.cpp:17-60)
    and the combined inverse mass at their shared nodes
    (Order2Newmark.cpp:32-66).  `fluid_vertex_rho` holds RHO at the four
    vertices of every fluid cell (the density rho_0 of AcousticToElastic2D is
    their mean).  Pair a fluid and a solid shot with Batch.couple."""

    def __init__(self, fluid, solid, fluid_vertex_rho):
        rho = np.ascontiguousarray(fluid_vertex_rho, dtype=np.float64).reshape(-1)
        if rho.size != 4 * fluid.E:
            raise SemError("fluid_vertex_rho needs 4 values per fluid cell")
        self.fluid, self.solid = fluid, solid
        self._h = _VP()
        _check(lib().sem_coupling_create(C.byref(self._h), fluid.handle, solid.handle, _d(rho)))
        info = np.zeros(4, dtype=np.int64)
        _check(lib().sem_coupling_info(self._h, _l(info)))
        self.shared_nodes, self.fluid_targets, self.solid_targets, self.terms = (int(x) for x in info)

    def close(self):
        if getattr(self, "_h", None):
            lib().sem_coupling_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h


def coupled_from_exodus(mesh_path, model_path=None, order=4, quirks=None, device=0):
    """The 2-D fluid-solid problem of an Exodus mesh and model whose elemental
    "fluid" flag assigns each element its physics, as the reference builds
    its coupled elements (ExodusModel::getElementType, ExodusModel.cpp:473-518;
    Mesh.cpp:314-333 and :480-521): returns (fluid Mesh, solid Mesh,
    Coupling).  The model defaults to the mesh file."""
    from salvus_amd import exodus, model
    dim, cells, coords, _ = exodus.read_mesh(mesh_path)
    if dim != 2:
        raise SemError("coupled_from_exodus: the fluid-solid coupling is 2-D (quads)")
    _, centres, variables = exodus.read_model(model_path or mesh_path)
    fluid, vp, solid_par, rho = model.coupled_parameters(cells, coords, centres, variables)
    if fluid.all() or not fluid.any():
        raise SemError("coupled_from_exodus: the model marks every element %s"
                       % ("fluid" if fluid.all() else "elastic"))
    (cf, xf, vp), (cs, xs, solid_par) = split_coupled(cells, coords, fluid, (vp, solid_par))
    q = DEFAULT_QUIRKS if quirks is None else quirks
    mf = Mesh(2, order, ACOUSTIC, cf, xf, vp, q, device)
    ms = Mesh(2, order, ELASTIC, cs, xs, solid_par, q, device)
    return mf, ms, Coupling(mf, ms, rho)


class Batch:
    """Shots advanced together (Problem::initializeElements source/receiver
    attach, Problem.cpp:34-142; the step loop of Main.cpp:45-66)."""

    def __init__(self, device=0):
        self._h = _VP()
        _check(lib().sem_batch_create(C.byref(self._h), device))
        self.meshes = []
        self.shots = []  # (mesh, nsteps, nrec)
        self.couplings = []  # kept alive while the batch refers to them

    def close(self):
        if getattr(self, "_h", None):
            lib().sem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_shot(self, mesh, nsteps, dt):
        k = C.c_int(0)
        _check(lib().sem_batch_add_shot(self._h, mesh.handle, nsteps, dt, C.byref(k)))
        self.meshes.append(mesh)
        self.shots.append([mesh, nsteps, 0])
        return k.value

    def couple(self, coupling, fluid_shot, solid_shot):
        """Couples a fluid and a solid shot of this batch (same dt and step
        count) through `coupling`; before prepare."""
        _check(lib().sem_batch_couple(self._h, coupling.handle, fluid_shot, solid_shot))
        self.couplings.append(coupling)

    def add_source(self, shot, loc, amp, f0, t0, direction=None):
        m = self.shots[shot][0]
        d = np.zeros(3)
        if direction is None:
            d[0] = 1.0
        else:
            d[: len(direction)] = direction
        _check(lib().sem_shot_add_source(self._h, shot, _d(np.asarray(loc, dtype=np.float64)),
                                         amp, f0, t0, _d(d)))

    def add_moment_source(self, shot, loc, M, amp, f0, t0):
        _check(lib().sem_shot_add_moment_source(
            self._h, shot, _d(np.asarray(loc, dtype=np.float64)),
            _d(np.ascontiguousarray(M, dtype=np.float64)), amp, f0, t0))

    def add_file_source(self, shot, loc, stf):
        """Tabulated point source: stf[n, k] is force component k at step n
        (SourceHdf5::fire row n); stf has one column per field component."""
        stf = np.asarray(stf, dtype=np.float64)
        if stf.ndim == 1:
            stf = stf[:, None]
        cols = np.ascontiguousarray(stf.T)  # component-major = column-major rows
        _check(lib().sem_shot_add_file_source(self._h, shot, _d(np.asarray(loc, dtype=np.float64)),
                                              stf.shape[1], stf.shape[0], _d(cols)))

    def add_receiver(self, shot, loc):
        _check(lib().sem_shot_add_receiver(self._h, shot, _d(np.asarray(loc, dtype=np.float64))))
        self.shots[shot][2] += 1

    def set_field(self, shot, which, vals):
        v = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
        _check(lib().sem_shot_set_field(self._h, shot, FIELD[which], _d(v)))

    def get_field(self, shot, which):
        m = self.shots[shot][0]
        out = np.zeros(m.c * m.G)
        _check(lib().sem_shot_get_field(self._h, shot, FIELD[which], _d(out)))
        return out.reshape(m.c, m.G)

    def prepare(self, path=PATH_FUSED, elems_per_block=0, shots_per_group=0):
        _check(lib().sem_batch_prepare(self._h, path, elems_per_block, shots_per_group))

    def run(self, nsteps):
        _check(lib().sem_batch_run(self._h, nsteps))

    def progress(self, shot):
        s, t = np.zeros(1, dtype=np.int64), np.zeros(1)
        _check(lib().sem_shot_progress(self._h, shot, _l(s), _d(t)))
        return int(s[0]), float(t[0])

    def seismograms(self, shot):
        m, nsteps, nrec = self.shots[shot]
        out = np.zeros(max(nsteps * nrec * m.c, 1), dtype=np.float32)
        _check(lib().sem_shot_get_seismograms(self._h, shot, _f(out)))
        return out[: nsteps * nrec * m.c].reshape(nsteps, nrec, m.c)

    def write_receivers_h5(self, shot, path, names, append=False):
        """Seismograms of one shot in the reference's receiver file format."""
        write_receivers_h5(path, names, self.seismograms(shot), append=append)

    def movie_frame(self, shot, path, names, append=True):
        """One movie frame of the named global fields of a shot (the
        reference's names: u, v, a, a_ or ux, uy, ..., ax_, ...) at the
        shot's current time (Problem::saveSolution, Problem.cpp:217-237)."""
        m = self.shots[shot][0]
        known = movie_field_names(m.c)
        for f in names:
            if f not in known:
                raise KeyError("You are attempting to save field %s which is not registered. "
                               "Registered fields are { %s }." % (f, " ".join(known)))
        got = {}
        frame = {}
        for f in names:
            which, k = known[f]
            if which not in got:
                got[which] = self.get_field(shot, [w for w, i in FIELD.items() if i == which][0])
            frame[f] = got[which][k]
        if not append:  # a new movie file starts with its mesh
            write_movie_mesh(path, m.dim, m.cells, m.coords, m.node_coords())
        write_movie_frame(path, frame, self.progress(shot)[1], True)

    def config(self):
        c = np.zeros(5, dtype=np.int64)
        _check(lib().sem_batch_config(self._h, _l(c)))
        return dict(elems_per_block=int(c[0]), shots_per_group=int(c[1]),
                    block_steps=int(c[2]), groups=int(c[3]), tiled_groups=int(c[4]))

    def lanes(self):
        n = C.c_int(0)
        _check(lib().sem_batch_lanes(self._h, C.byref(n)))
        return n.value

    def kernel_time(self):
        ms, n = np.zeros(1), np.zeros(1, dtype=np.int64)
        _check(lib().sem_batch_kernel_time(self._h, _d(ms), _l(n)))
        return float(ms[0]), int(n[0])


class Order2Newmark:
    """One shot driven like the reference's Problem/Order2Newmark loop
    (Main.cpp:45-66): explicit Newmark with the reference's update order, the
    stiffness in FP32 on the device and Kahan-compensated FP32 u/v."""

    def __init__(self, mesh, dt, nsteps, path=PATH_FUSED, device=0):
        self.mesh, self.dt, self.nsteps, self.path = mesh, dt, nsteps, path
        self.batch = Batch(device)
        self.shot = self.batch.add_shot(mesh, nsteps, dt)
        self._prepared = False

    def add_source(self, loc, amp, f0, t0, direction=None):
        self.batch.add_source(self.shot, loc, amp, f0, t0, direction)

    def add_moment_source(self, loc, M, amp, f0, t0):
        self.batch.add_moment_source(self.shot, loc, M, amp, f0, t0)

    def add_file_source(self, loc, stf):
        self.batch.add_file_source(self.shot, loc, stf)

    def add_receiver(self, loc):
        self.batch.add_receiver(self.shot, loc)

    def set_field(self, which, vals):
        self.batch.set_field(self.shot, which, vals)

    def get_field(self, which):
        return self.batch.get_field(self.shot, which)

    def initializeGlobalDofs(self):
        if not self._prepared:
            self.batch.prepare(self.path)
            self._prepared = True

    def takeTimeStep(self, nsteps=1):
        self.initializeGlobalDofs()
        self.batch.run(nsteps)

    def seismograms(self):
        return self.batch.seismograms(self.shot)

    def time(self):
        return self.batch.progress(self.shot)[1]

    def saveSolution(self, path, save_fields, append=True):
        """Problem::saveSolution (Problem.cpp:217-237): one movie frame of
        save_fields at the current time; the driver calls it after every
        step (Main.cpp:55-62)."""
        self.initializeGlobalDofs()
        self.batch.movie_frame(self.shot, path, save_fields, append)


def source_record(src):
    """One SOURCE_DOUBLES record: kind (0 force, 1 moment tensor), location,
    amplitude, Ricker f0 and t0, force direction, moment tensor (row-major)."""
    rec = np.zeros(SOURCE_DOUBLES)
    moment = src.get("kind", "force") == "moment"
    rec[0] = 1.0 if moment else 0.0
    loc = np.asarray(src["loc"], dtype=np.float64)
    rec[1:1 + len(loc)] = loc
    rec[4], rec[5], rec[6] = src.get("amp", 1.0), src["f0"], src["t0"]
    if moment:
        M = np.asarray(src["M"], dtype=np.float64)
        d = M.shape[0]
        for i in range(d):
            for j in range(d):
                rec[10 + 3 * i + j] = M[i, j]
    else:
        d = np.asarray(src.get("dir", (1.0,)), dtype=np.float64)
        rec[7:7 + len(d)] = d
    return rec


class Server:
    """Continuous-batching request loop over device-resident meshes: each
    request is one shot; active shots advance together block_steps steps per
    pump(); finished ones are retired and arrived ones admitted in arrival
    order while the HBM budget (bytes, 0 = free memory less a reserve) and
    max_active (0 = no cap) allow.  On the fused path the head of the queue
    is admitted with up to shots_per_group - 1 later arrived requests on its
    mesh and dt whose step counts lie within step_spread of the largest (one
    multi-shot group)."""

    STATES = {0: "pending", 1: "active", 2: "done", 3: "failed", 4: "cancelled"}

    def __init__(self, path=PATH_FUSED, hbm_budget=0, max_active=0, block_steps=32, device=0,
                 shots_per_group=4, step_spread=0.1):
        self._h = _VP()
        _check(lib().sem_server_create(C.byref(self._h), device, path, int(hbm_budget),
                                       max_active, block_steps))
        _check(lib().sem_server_set_grouping(self._h, int(shots_per_group), float(step_spread)))
        self.requests = []  # (mesh, nsteps, nrec)

    def close(self):
        if getattr(self, "_h", None):
            lib().sem_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def clock(self):
        t = np.zeros(1)
        _check(lib().sem_server_clock(self._h, _d(t)))
        return float(t[0])

    def submit(self, mesh, nsteps, dt, sources, receivers, arrival=0.0):
        src = np.ascontiguousarray(np.array([source_record(s) for s in sources]).reshape(-1))
        if src.size == 0:
            src = np.zeros(SOURCE_DOUBLES)
        rec = np.zeros((max(len(receivers), 1), 3))
        for k, r in enumerate(receivers):
            rec[k, :len(r)] = r
        rec = np.ascontiguousarray(rec.reshape(-1))
        k = C.c_int(0)
        _check(lib().sem_server_submit(self._h, mesh.handle, int(nsteps), float(dt), float(arrival),
                                       len(sources), _d(src), len(receivers), _d(rec),
                                       C.byref(k)))
        self.requests.append((mesh, int(nsteps), len(receivers)))
        return k.value

    def pump(self):
        st = np.zeros(6)
        _check(lib().sem_server_pump(self._h, _d(st)))
        return dict(active=int(st[0]), pending=int(st[1]), done=int(st[2]), steps=int(st[3]),
                    next_arrival=float(st[4]), bytes=int(st[5]))

    def request(self, rid):
        info = np.zeros(6)
        _check(lib().sem_server_request(self._h, rid, _d(info)))
        shots = C.c_int(0)
        _check(lib().sem_server_request_group(self._h, rid, C.byref(shots)))
        return dict(state=self.STATES[int(info[0])], arrival=info[1], admit=info[2], done=info[3],
                    steps=int(info[4]), bytes=int(info[5]), group_shots=shots.value)

    def cancel(self, rid):
        """A pending request leaves the queue; an active one is retired at
        once (partial results dropped; device memory released with its
        group)."""
        _check(lib().sem_server_cancel(self._h, rid))

    def result(self, rid):
        mesh, nsteps, nrec = self.requests[rid]
        out = np.zeros(max(nsteps * nrec * mesh.c, 1), dtype=np.float32)
        _check(lib().sem_server_result(self._h, rid, _f(out)))
        return out[: nsteps * nrec * mesh.c].reshape(nsteps, nrec, mesh.c)

    def drain(self, idle_sleep=0.0005, progress=None, every=10.0):
        """Pumps until every submitted request is finished; progress(status)
        is called about every `every` seconds of the server clock."""
        import time
        last = self.clock()
        while True:
            st = self.pump()
            if progress is not None and self.clock() - last >= every:
                last = self.clock()
                progress(st)
            if st["active"] == 0 and st["pending"] == 0:
                return
            if st["active"] == 0 and st["next_arrival"] > 0:
                time.sleep(max(0.0, min(idle_sleep, st["next_arrival"] - self.clock())))
