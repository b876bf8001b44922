"""GPU tests of sem_mesh_apply_stiffness_noalloc, the allocation-free variant
of sem_mesh_apply_stiffness (device buffers in and out, descriptors in the
library's static device scratch).  Called through the C ABI with device
buffers from the HIP runtime; checked against the allocating entry point
(bit for bit where both take the element kernel) and against the FP64
oracle, on consecutive calls of which the second is smaller."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import sem
from tests import problems as pb

pytestmark = pytest.mark.gpu

TOL_OP = 2e-6


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def get(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0


def noalloc(M, u, hip, stream=None):
    L = sem.lib()
    L.sem_mesh_apply_stiffness_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    u32 = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
    du, dk = hip.alloc(u32.nbytes), hip.alloc(u32.nbytes)
    try:
        hip.put(du, u32)
        hip.put(dk, np.full_like(u32, np.nan))  # the output starts as garbage
        rc = L.sem_mesh_apply_stiffness_noalloc(M._h, du, dk, stream)
        if rc != 0:
            raise sem.SemError(L.sem_last_error().decode())
        if stream is not None:
            assert hip.L.hipStreamSynchronize(stream) == 0
        out = np.empty_like(u32)
        hip.get(out, dk)
    finally:
        hip.L.hipFree(du)
        hip.L.hipFree(dk)
    return out.reshape(M.c, M.G)


def field(S, dim, seed):
    rng = np.random.default_rng(seed)
    xyz = S.node_coords()
    return np.stack([np.sin(0.05 * xyz @ rng.normal(size=dim)) + 0.1 * rng.normal(size=S.G)
                     for _ in range(S.c)])


# (dim, physics, elements per side, distorted): the first call larger than
# the second; distorted meshes take the element kernel in both entry points
CALLS = [[(3, 1, 5, True), (2, 0, 6, True)],
         [(3, 0, 6, False), (3, 1, 3, False)],
         [(2, 1, 9, True), (2, 1, 4, False)]]


@pytest.mark.parametrize("calls", CALLS)
def test_noalloc_matches_allocating_entry(calls):
    hip = Hip()
    stream = C.c_void_p()
    assert hip.L.hipStreamCreate(C.byref(stream)) == 0
    try:
        sizes = []
        for k, (dim, physics, nel, distorted) in enumerate(calls):
            P = pb.make(dim, physics, nel, order=4, seed=20 + k, distorted=distorted)
            M = sem.Mesh(P["dim"], 4, P["physics"], P["cells"], P["coords"], P["par"])
            S = po.Sim(P["dim"], 4, P["physics"], P["cells"], P["coords"], P["par"])
            u = field(S, dim, 30 + k)
            ref = S.apply_stiffness(u)
            got = noalloc(M, u, hip, stream if k == 1 else None)
            assert np.all(np.isfinite(got))
            assert pb.rel_l2(got, ref) < TOL_OP
            alloc = M.apply_stiffness(u, sem.PATH_FUSED)
            if distorted:
                np.testing.assert_array_equal(got, alloc)  # same kernel, same order
            else:
                assert pb.rel_l2(got, alloc) < TOL_OP  # tile kernel vs element kernel
            sizes.append(M.c * M.G)
        assert sizes[1] < sizes[0]
    finally:
        hip.L.hipStreamDestroy(stream)
