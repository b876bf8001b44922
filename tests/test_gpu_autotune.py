"""Autotuners on the GPU: the tile shape (elements per workgroup) a mesh picks
is a compiled one and keeps oracle parity; the shots-per-group choice of a
batch changes no result bit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import sem
from tests import problems as pb

pytestmark = pytest.mark.gpu

TOL_RUN = 1e-5
COMPILED_3D = {(4, 2, 2), (2, 2, 2), (4, 4, 2)}


@pytest.mark.parametrize("physics,shots", [(0, 1), (1, 1), (1, 3)])
def test_mesh_autotune_picks_compiled_shape(physics, shots):
    P = pb.make(3, physics, 5, order=4, seed=40 + physics, nsteps=100)
    M = sem.Mesh(3, 4, physics, P["cells"], P["coords"], P["par"])
    shape = M.autotune(reps=2, shots=shots)
    assert shape in COMPILED_3D and M.ntiles > 0
    O = po.Sim(3, 4, physics, P["cells"], P["coords"], P["par"])
    s = P["src"]
    O.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in P["recs"]:
        O.add_receiver(r)
    ref = O.run(P["dt"], P["nsteps"])
    prob = sem.Order2Newmark(M, P["dt"], P["nsteps"])
    prob.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in P["recs"]:
        prob.add_receiver(r)
    prob.takeTimeStep(P["nsteps"])
    assert pb.rel_l2(prob.seismograms(), ref) < TOL_RUN


def test_shots_per_group_autotune_changes_no_bit():
    P = pb.make(3, 1, 4, order=4, seed=44, nsteps=60)
    M = sem.Mesh(3, 4, 1, P["cells"], P["coords"], P["par"])
    s = P["src"]
    scales = (1.0, 0.9, 1.1, 0.95, 1.05)

    def run(spg):
        B = sem.Batch()
        ids = []
        for f in scales:
            k = B.add_shot(M, P["nsteps"], P["dt"])
            B.add_source(k, np.asarray(s["loc"]) * f, s["amp"], s["f0"], s["t0"], s["dir"])
            for r in P["recs"]:
                B.add_receiver(k, r)
            ids.append(k)
        B.prepare(sem.PATH_FUSED, 0, spg)
        B.run(P["nsteps"])
        return B.config()["shots_per_group"], [B.seismograms(k) for k in ids]

    chosen, tuned = run(0)
    assert chosen in (1, 2, 4)
    one, single = run(1)
    assert one == 1
    for a, b in zip(tuned, single):
        np.testing.assert_array_equal(a, b)
