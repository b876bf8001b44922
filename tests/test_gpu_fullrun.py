"""The accuracy gate at the configurations' full length: both FP32 HIP paths
against FP64 oracle runs of the whole workload length, committed as
fixtures by tests/golden/make_fullrun_fixtures.py:

  cfg2     2-D elastic at its full size, 500^2 quads, all 5000 steps
  cfg3-32  the 3-D acoustic workload's 32^3 sub-box, all 2000 steps
  cfg4-32  the 3-D elastic workload's 32^3 mesh, two moment-tensor shots,
           all 1500 steps each

Bars (the metric's): receiver seismograms relative L2 <= 1e-5 over every
receiver, component and step; the final displacement relative L2 <= 1e-5
on the fixture's strided sample, and its L2 norm within 1e-5.  The inputs
are regenerated from the recorded seeds and checked against the fixture's
SHA-256 of cells, coordinates and material first."""
import ast
from pathlib import Path

import numpy as np
import pytest

from salvus_amd import sem
from tests import problems as pb
from tests.golden.make_fullrun_fixtures import input_digest, problem

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLDEN = Path(__file__).resolve().parent / "golden"


def add_all(B, k, sources, receivers):
    for s in sources:
        if s["kind"] == "moment":
            B.add_moment_source(k, s["loc"], s["M"], s["amp"], s["f0"], s["t0"])
        else:
            B.add_source(k, s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in receivers:
        B.add_receiver(k, r)


@pytest.mark.parametrize("name", ["cfg3-32", "cfg4-32", "cfg2"])
def test_full_length_against_fp64(name):
    F = np.load(GOLDEN / ("fullrun_%s.npz" % name))
    dim, phys, cells, coords, par, dt, nsteps, shots, recs, seeds = problem(name)
    assert repr(seeds) == str(F["seeds"])
    assert input_digest(cells, coords, par) == str(F["digest"]), "regenerated inputs differ"
    assert nsteps == int(F["nsteps"]) and dt == float(F["dt"])
    stride = int(F["stride"])
    M = sem.Mesh(dim, 4, phys, cells, coords, par)
    for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
        B = sem.Batch()
        ks = []
        for sources in shots:
            k = B.add_shot(M, nsteps, dt)
            add_all(B, k, sources, recs)
            ks.append(k)
        B.prepare(path)
        B.run(nsteps)
        for i, k in enumerate(ks):
            ref = F["seis%d" % i].astype(np.float64)
            e_r = pb.rel_l2(B.seismograms(k), ref)
            u = B.get_field(k, "u").ravel()
            e_u = pb.rel_l2(u[::stride], F["u_sample%d" % i])
            e_n = abs(np.linalg.norm(u) / float(F["u_norm%d" % i]) - 1.0)
            print("%s path %d shot %d: receivers %.3g, final u (sample) %.3g, |u| %.3g"
                  % (name, path, i, e_r, e_u, e_n), flush=True)
            assert e_r <= TOL and e_u <= TOL and e_n <= TOL
        B.close()
    M.close()
