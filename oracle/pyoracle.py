"""ctypes front end of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker.  The product
(salvus_amd) never imports this module.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_LIB_PATH = HERE / "_ref" / "libgll_ref.so"

ACOUSTIC, ELASTIC = 0, 1
QUIRK_NEWMARK3D, QUIRK_E3D_SOURCE, QUIRK_QUAD_JACOBIAN = 1, 2, 4
DEFAULT_QUIRKS = QUIRK_E3D_SOURCE | QUIRK_QUAD_JACOBIAN

_lib = None

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)


def build(force=False):
    """Compile liboracle.so (and the reference GLL table library when the
    reference sources are present); force recompiles everything."""
    import subprocess
    flags = ["-B"] if force else []
    subprocess.run(["make", "-s"] + flags + ["-C", str(HERE)], check=True)
    if Path("/root/reference/src").is_dir():
        subprocess.run(["make", "-s"] + flags + ["-C", str(HERE), "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_plex_new.restype = C.c_void_p
        L.orc_plex_new.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, I32]
        L.orc_plex_free.argtypes = [C.c_void_p]
        for f in ("orc_plex_counts",):
            getattr(L, f).argtypes = [C.c_void_p, I64]
        for f in ("orc_plex_conn", "orc_plex_cell_cones", "orc_plex_edge_cones",
                  "orc_plex_face_cones"):
            getattr(L, f).argtypes = [C.c_void_p, I32]
        L.orc_plex_offsets.argtypes = [C.c_void_p, I64]
        L.orc_color_elements.argtypes = [C.c_int, C.c_int64, C.c_int64, I32, I32]
        L.orc_sim_new.restype = C.c_void_p
        L.orc_sim_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, I32, D, D,
                                  C.c_int]
        L.orc_sim_free.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_sim_nglobal.restype = C.c_int64
        L.orc_sim_nglobal.argtypes = [C.c_void_p]
        L.orc_sim_ncomp.argtypes = [C.c_void_p]
        L.orc_sim_conn.argtypes = [C.c_void_p, I32]
        L.orc_sim_inverse_mass.argtypes = [C.c_void_p, D]
        L.orc_sim_node_coords.argtypes = [C.c_void_p, D]
        L.orc_sim_set_dirichlet.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.orc_sim_add_source.argtypes = [C.c_void_p, D, C.c_double, C.c_double, C.c_double, D]
        L.orc_sim_add_moment_source.argtypes = [C.c_void_p, D, D, C.c_double, C.c_double,
                                                C.c_double]
        L.orc_sim_add_receiver.argtypes = [C.c_void_p, D]
        L.orc_sim_source_info.argtypes = [C.c_void_p, C.c_int, I64, D]
        L.orc_sim_receiver_info.argtypes = [C.c_void_p, C.c_int, I64, D]
        L.orc_sim_source_load.argtypes = [C.c_void_p, C.c_int, D]
        L.orc_sim_receiver_weights.argtypes = [C.c_void_p, C.c_int, D]
        L.orc_sim_set_field.argtypes = [C.c_void_p, C.c_int, D]
        L.orc_sim_get_field.argtypes = [C.c_void_p, C.c_int, D]
        L.orc_sim_run.argtypes = [C.c_void_p, C.c_double, C.c_int64, D, C.c_int]
        L.orc_sim_time.restype = C.c_double
        L.orc_sim_time.argtypes = [C.c_void_p]
        L.orc_sim_apply_stiffness.argtypes = [C.c_void_p, D, D]
        L.orc_reference_step_count.restype = C.c_int64
        L.orc_reference_step_count.argtypes = [C.c_double, C.c_double, D]
        L.orc_ricker.restype = C.c_double
        L.orc_ricker.argtypes = [C.c_double] * 4
        L.orc_gll.argtypes = [C.c_int, D, D]
        L.orc_gll_exact.argtypes = [C.c_int, D, D]
        L.orc_deriv_matrix.argtypes = [C.c_int, D]
        L.orc_lagrange.argtypes = [C.c_int, D, C.c_double, D]
        L.orc_lagrange_deriv.argtypes = [C.c_int, D, C.c_double, D]
        L.orc_quad_invjac.argtypes = [C.c_double, C.c_double, D, D, D]
        L.orc_quad_invjac_consistent.argtypes = [C.c_double, C.c_double, D, D, D]
        L.orc_quad_interp.argtypes = [C.c_double, C.c_double, D]
        L.orc_quad_check_hull.argtypes = [C.c_double, C.c_double, D]
        L.orc_quad_inverse_map.argtypes = [C.c_double, C.c_double, D, D]
        L.orc_quad_nodal_points.argtypes = [C.c_int, D, D]
        L.orc_hex_invjac.argtypes = [C.c_double] * 3 + [D, D, D]
        L.orc_hex_interp.argtypes = [C.c_double] * 3 + [D]
        L.orc_hex_check_hull.argtypes = [C.c_double] * 3 + [D]
        L.orc_hex_inverse_map.argtypes = [C.c_double] * 3 + [D, D]
        L.orc_hex_nodal_points.argtypes = [C.c_int, D, D]
        for f in ("orc_elem_gradient", "orc_elem_grad_test_integrate", "orc_elem_test_integrate",
                  "orc_elem_delta_coeffs"):
            getattr(L, f).argtypes = [C.c_int, C.c_int, D, D, D]
        L.orc_elem_par_at_int_pts.argtypes = [C.c_int, C.c_int, D, D]
        L.orc_elem_stiffness.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own GLL tables compiled from its sources (or None)."""
    if not REF_LIB_PATH.exists():
        return None
    return C.CDLL(str(REF_LIB_PATH))


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(I32)


def lp(a):
    return a.ctypes.data_as(I64)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- GLL
def gll(order):
    xi, w = np.zeros(order + 1), np.zeros(order + 1)
    if lib().orc_gll(order, dp(xi), dp(w)):
        raise ValueError("order")
    return xi, w


def gll_exact(order):
    xi, w = np.zeros(order + 1), np.zeros(order + 1)
    lib().orc_gll_exact(order, dp(xi), dp(w))
    return xi, w


def deriv_matrix(order):
    Dm = np.zeros((order + 1) * (order + 1))
    lib().orc_deriv_matrix(order, dp(Dm))
    return Dm.reshape(order + 1, order + 1)


def lagrange(xi, x):
    xi = f64(xi)
    out = np.zeros(len(xi))
    lib().orc_lagrange(len(xi), dp(xi), x, dp(out))
    return out


# ---------------------------------------------------------------- geometry
def quad_invjac(r, s, vtx):
    v = f64(vtx).reshape(-1)
    det, inv = np.zeros(1), np.zeros(4)
    lib().orc_quad_invjac(r, s, dp(v), dp(det), dp(inv))
    return det[0], inv.reshape(2, 2)


def quad_invjac_consistent(r, s, vtx):
    v = f64(vtx).reshape(-1)
    det, inv = np.zeros(1), np.zeros(4)
    lib().orc_quad_invjac_consistent(r, s, dp(v), dp(det), dp(inv))
    return det[0], inv.reshape(2, 2)


def hex_invjac(r, s, t, vtx):
    v = f64(vtx).reshape(-1)
    det, inv = np.zeros(1), np.zeros(9)
    lib().orc_hex_invjac(r, s, t, dp(v), dp(det), dp(inv))
    return det[0], inv.reshape(3, 3)


def check_hull(pt, vtx):
    v = f64(vtx).reshape(-1)
    if len(pt) == 2:
        return lib().orc_quad_check_hull(pt[0], pt[1], dp(v))
    return lib().orc_hex_check_hull(pt[0], pt[1], pt[2], dp(v))


def inverse_map(pt, vtx):
    v = f64(vtx).reshape(-1)
    out = np.zeros(len(pt))
    if len(pt) == 2:
        rc = lib().orc_quad_inverse_map(pt[0], pt[1], dp(v), dp(out))
    else:
        rc = lib().orc_hex_inverse_map(pt[0], pt[1], pt[2], dp(v), dp(out))
    if rc:
        raise RuntimeError("inverseCoordinateTransform failed to converge")
    return out


def interp(ref):
    out = np.zeros(4 if len(ref) == 2 else 8)
    if len(ref) == 2:
        lib().orc_quad_interp(ref[0], ref[1], dp(out))
    else:
        lib().orc_hex_interp(ref[0], ref[1], ref[2], dp(out))
    return out


def nodal_points(order, vtx):
    v = f64(vtx).reshape(-1)
    dim = 2 if v.size == 8 else 3
    out = np.zeros((order + 1) ** dim * dim)
    (lib().orc_quad_nodal_points if dim == 2 else lib().orc_hex_nodal_points)(order, dp(v), dp(out))
    return out.reshape(-1, dim)


# ---------------------------------------------------------------- element ops
def _elem(fn, dim, order, vtx, arr, out_size):
    v = f64(vtx).reshape(-1)
    a = f64(arr).reshape(-1)
    out = np.zeros(out_size)
    if getattr(lib(), fn)(dim, order, dp(v), dp(a), dp(out)):
        raise ValueError(fn)
    return out


def elem_gradient(dim, order, vtx, u):
    nd = (order + 1) ** dim
    return _elem("orc_elem_gradient", dim, order, vtx, u, nd * dim).reshape(nd, dim)


def elem_grad_test_integrate(dim, order, vtx, f):
    return _elem("orc_elem_grad_test_integrate", dim, order, vtx, f, (order + 1) ** dim)


def elem_test_integrate(dim, order, vtx, f):
    return _elem("orc_elem_test_integrate", dim, order, vtx, f, (order + 1) ** dim)


def elem_delta_coeffs(dim, order, vtx, ref):
    return _elem("orc_elem_delta_coeffs", dim, order, vtx, ref, (order + 1) ** dim)


def elem_par_at_int_pts(dim, order, vtx_par):
    a = f64(vtx_par)
    out = np.zeros((order + 1) ** dim)
    lib().orc_elem_par_at_int_pts(dim, order, dp(a), dp(out))
    return out


def elem_stiffness(dim, order, physics, vtx, par, u):
    nd = (order + 1) ** dim
    c = 1 if physics == ACOUSTIC else dim
    v, p, uu = f64(vtx).reshape(-1), f64(par).reshape(-1), f64(u).reshape(-1)
    out = np.zeros(nd * c)
    lib().orc_elem_stiffness(dim, order, physics, dp(v), dp(p), dp(uu), dp(out))
    return out


# ---------------------------------------------------------------- plex-lite
class Plex:
    def __init__(self, dim, order, cells, nvert):
        cells = i32(cells)
        self.dim, self.order = dim, order
        nv = 4 if dim == 2 else 8
        self.E = cells.size // nv
        self._p = lib().orc_plex_new(dim, order, self.E, nvert, ip(cells))
        if not self._p:
            raise ValueError("plex")
        cnt = np.zeros(6, dtype=np.int64)
        lib().orc_plex_counts(self._p, lp(cnt))
        self.E, self.V, self.NF, self.NEd, self.npoints, self.nglobal = (int(x) for x in cnt)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_plex_free(self._p)
            self._p = None

    def conn(self):
        nd = (self.order + 1) ** self.dim
        out = np.zeros(self.E * nd, dtype=np.int32)
        lib().orc_plex_conn(self._p, ip(out))
        return out.reshape(self.E, nd)

    def cell_cones(self):
        k = 4 if self.dim == 2 else 6
        out = np.zeros(self.E * k, dtype=np.int32)
        lib().orc_plex_cell_cones(self._p, ip(out))
        return out.reshape(self.E, k)

    def edge_cones(self):
        out = np.zeros(self.NEd * 2, dtype=np.int32)
        lib().orc_plex_edge_cones(self._p, ip(out))
        return out.reshape(-1, 2)

    def face_cones(self):
        out = np.zeros(max(self.NF, 0) * 4, dtype=np.int32)
        if self.NF:
            lib().orc_plex_face_cones(self._p, ip(out))
        return out.reshape(-1, 4)

    def offsets(self):
        out = np.zeros(self.npoints, dtype=np.int64)
        lib().orc_plex_offsets(self._p, lp(out))
        return out


def color_elements(dim, cells, nvert):
    cells = i32(cells)
    nv = 4 if dim == 2 else 8
    E = cells.size // nv
    col = np.zeros(E, dtype=np.int32)
    n = lib().orc_color_elements(dim, E, nvert, ip(cells), ip(col))
    return col, n


# ---------------------------------------------------------------- simulation
class Sim:
    """FP64 restatement of the reference's Order2Newmark problem."""

    def __init__(self, dim, order, physics, cells, coords, par, quirks=DEFAULT_QUIRKS):
        cells, coords, par = i32(cells), f64(coords), f64(par)
        nv = 4 if dim == 2 else 8
        self.dim, self.order, self.physics = dim, order, physics
        self.E = cells.size // nv
        self.V = coords.size // dim
        self._s = lib().orc_sim_new(dim, order, physics, self.E, self.V, ip(cells), dp(coords),
                                    dp(par), quirks)
        if not self._s:
            raise RuntimeError(lib().orc_last_error().decode())
        self.G = int(lib().orc_sim_nglobal(self._s))
        self.c = int(lib().orc_sim_ncomp(self._s))
        self.nrec = 0
        self.nsrc = 0

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_sim_free(self._s)
            self._s = None

    def conn(self):
        nd = (self.order + 1) ** self.dim
        out = np.zeros(self.E * nd, dtype=np.int32)
        lib().orc_sim_conn(self._s, ip(out))
        return out.reshape(self.E, nd)

    def inverse_mass(self):
        out = np.zeros(self.G)
        lib().orc_sim_inverse_mass(self._s, dp(out))
        return out

    def node_coords(self):
        out = np.zeros(self.G * self.dim)
        lib().orc_sim_node_coords(self._s, dp(out))
        return out.reshape(self.G, self.dim)

    def set_dirichlet(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        lib().orc_sim_set_dirichlet(self._s, m.ctypes.data_as(C.POINTER(C.c_uint8)))

    def add_source(self, loc, amp, f0, t0, direction=None):
        d = f64(direction if direction is not None else [1.0] + [0.0] * (self.c - 1))
        rc = lib().orc_sim_add_source(self._s, dp(f64(loc)), amp, f0, t0, dp(d))
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        self.nsrc += 1
        return rc

    def add_moment_source(self, loc, M, amp, f0, t0):
        rc = lib().orc_sim_add_moment_source(self._s, dp(f64(loc)), dp(f64(M)), amp, f0, t0)
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        self.nsrc += 1
        return rc

    def add_receiver(self, loc):
        rc = lib().orc_sim_add_receiver(self._s, dp(f64(loc)))
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        self.nrec += 1
        return rc

    def source_info(self, k):
        e, ref = np.zeros(1, dtype=np.int64), np.zeros(3)
        lib().orc_sim_source_info(self._s, k, lp(e), dp(ref))
        return int(e[0]), ref[: self.dim]

    def receiver_info(self, k):
        e, ref = np.zeros(1, dtype=np.int64), np.zeros(3)
        lib().orc_sim_receiver_info(self._s, k, lp(e), dp(ref))
        return int(e[0]), ref[: self.dim]

    def source_load(self, k):
        nd = (self.order + 1) ** self.dim
        out = np.zeros(nd * self.c)
        lib().orc_sim_source_load(self._s, k, dp(out))
        return out.reshape(self.c, nd)

    def receiver_weights(self, k):
        out = np.zeros((self.order + 1) ** self.dim)
        lib().orc_sim_receiver_weights(self._s, k, dp(out))
        return out

    def set_field(self, which, vals):
        lib().orc_sim_set_field(self._s, {"u": 0, "v": 1, "a_": 2}[which],
                                dp(f64(vals).reshape(-1)))

    def get_field(self, which):
        out = np.zeros(self.c * self.G)
        lib().orc_sim_get_field(self._s, {"u": 0, "v": 1, "a_": 2, "a": 3}[which], dp(out))
        return out.reshape(self.c, self.G)

    def run(self, dt, nsteps, nthreads=1):
        seis = np.zeros(max(nsteps * self.nrec * self.c, 1))
        lib().orc_sim_run(self._s, dt, nsteps, dp(seis), nthreads)
        return seis[: nsteps * self.nrec * self.c].reshape(nsteps, self.nrec, self.c)

    def time(self):
        return lib().orc_sim_time(self._s)

    def apply_stiffness(self, u):
        u = f64(u).reshape(-1)
        out = np.zeros(self.c * self.G)
        lib().orc_sim_apply_stiffness(self._s, dp(u), dp(out))
        return out.reshape(self.c, self.G)


def reference_step_count(duration, dt):
    dto = np.zeros(1)
    n = lib().orc_reference_step_count(duration, dt, dp(dto))
    return int(n), float(dto[0])


def ricker(t, amp, f0, t0):
    return lib().orc_ricker(t, amp, f0, t0)
