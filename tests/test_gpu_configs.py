"""Parity on the benchmark configurations: the reference's own CPU-runnable
case (cfg1: 2-D acoustic 50x50 quads, order 4, 1000 steps) at full size, and
the other configurations' laws (material, source type, receiver layout) on
sub-boxes the FP64 oracle runs in seconds (cfg2: 100^2 elastic quads; cfg3:
16^3 acoustic hexes, and the 32^3 sub-box for 200 steps; cfg4: 10^3 elastic
hexes with a moment-tensor source).
Both FP32 HIP paths must reach receiver and final-wavefield relative L2 of
1e-5 against the oracle."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import sem, workloads
from tests import problems as pb

pytestmark = pytest.mark.gpu

TOL = 1e-5
THREADS = min(os.cpu_count() or 1, 16)


def sample(name):
    if name == "cfg1":
        W = workloads.cfg1()
        return dict(dim=2, physics=0, cells=W["cells"], coords=W["coords"], par=W["par"],
                    dt=W["dt"], nsteps=W["nsteps"], sources=W["sources"], receivers=W["receivers"])
    if name == "cfg4":
        W = workloads.cfg4(sizes=(10,), shots_per_mesh=1, nsteps=250)
        mw = W["meshes"][0]
        return dict(dim=3, physics=1, cells=mw["cells"], coords=mw["coords"], par=mw["par"],
                    dt=mw["dt"], nsteps=250, sources=mw["shots"], receivers=mw["receivers"])
    if name == "cfg3-32":
        W = workloads.cfg3(n=32, nsteps=200, nrec_side=8)
    elif name == "cfg3":
        W = workloads.cfg3(n=16, nsteps=300, nrec_side=4)
    else:
        W = workloads.cfg2(n=100, nsteps=300)
    return dict(dim=W["dim"], physics=W["physics"], cells=W["cells"], coords=W["coords"],
                par=W["par"], dt=W["dt"], nsteps=W["nsteps"], sources=W["sources"],
                receivers=W["receivers"])


def add_all(target, sm, shot=None):
    args = () if shot is None else (shot,)
    for s in sm["sources"]:
        if s["kind"] == "moment":
            target.add_moment_source(*args, s["loc"], s["M"], s["amp"], s["f0"], s["t0"])
        else:
            target.add_source(*args, s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in sm["receivers"]:
        target.add_receiver(*args, r)


@pytest.mark.parametrize("name", ["cfg1", "cfg2", "cfg3", "cfg3-32", "cfg4"])
def test_config_parity(name):
    sm = sample(name)
    O = po.Sim(sm["dim"], 4, sm["physics"], sm["cells"], sm["coords"], sm["par"])
    add_all(O, sm)
    ref = O.run(sm["dt"], sm["nsteps"], nthreads=THREADS)
    ref_u = O.get_field("u")
    assert np.abs(ref).max() > 0
    M = sem.Mesh(sm["dim"], 4, sm["physics"], sm["cells"], sm["coords"], sm["par"])
    for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
        B = sem.Batch()
        k = B.add_shot(M, sm["nsteps"], sm["dt"])
        add_all(B, sm, k)
        B.prepare(path)
        B.run(sm["nsteps"])
        err_r = pb.rel_l2(B.seismograms(k), ref)
        err_u = pb.rel_l2(B.get_field(k, "u"), ref_u)
        print("%s path %d: receivers %.3g, final u %.3g" % (name, path, err_r, err_u), flush=True)
        assert err_r < TOL and err_u < TOL
        B.close()
