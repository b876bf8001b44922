"""The C ABI driven from C++ exactly as the reference-side binding drives it
(tests/cxx/hip_problem_driver.cpp: status codes turned into
std::runtime_error, one sem_batch_run of the whole loop): its seismograms
and final displacement equal the Python (ctypes) path's bit for bit and the
FP64 oracle's within the run tolerance, and an error (a receiver outside the
mesh) arrives as an exception."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import sem
from tests import problems as pb

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "cxx" / "hip_problem_driver"


def write_problem(path, P):
    dim, physics = P["dim"], P["physics"]
    cells = np.ascontiguousarray(P["cells"], dtype=np.int32)
    coords = np.ascontiguousarray(P["coords"], dtype=np.float64)
    par = np.ascontiguousarray(P["par"], dtype=np.float64).reshape(-1)
    recs = np.zeros((len(P["recs"]), 3))
    for k, r in enumerate(P["recs"]):
        recs[k, :dim] = r
    src = P["src"]
    loc, direction = np.zeros(3), np.zeros(3)
    loc[:dim] = src["loc"]
    direction[:len(src["dir"])] = src["dir"]
    with open(path, "wb") as f:
        f.write(struct.pack("6i", dim, P["order"], physics, len(cells), len(coords), len(recs)))
        f.write(struct.pack("q", P["nsteps"]))
        f.write(np.array([P["dt"], *loc, src["amp"], src["f0"], src["t0"], *direction]).tobytes())
        f.write(cells.tobytes())
        f.write(coords.tobytes())
        f.write(par.tobytes())
        f.write(recs.tobytes())


@pytest.mark.parametrize("dim,physics", [(2, 1), (3, 0), (3, 1)])
def test_cxx_driver_matches_ctypes_path_and_oracle(dim, physics, tmp_path):
    assert DRIVER.exists(), "build the driver first (__graft_entry__.build())"
    P = pb.make(dim, physics, 8 if dim == 2 else 4, order=4, seed=40 + dim, nsteps=150)
    write_problem(tmp_path / "p.bin", P)
    run = subprocess.run([str(DRIVER), str(tmp_path / "p.bin"), str(tmp_path / "o.bin")],
                         capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr)
    assert run.returncode == 0
    assert "error path: " in run.stdout and "outside" in run.stdout
    c, nrec = P["c"], len(P["recs"])
    raw = (tmp_path / "o.bin").read_bytes()
    ns = P["nsteps"] * nrec * c
    seis = np.frombuffer(raw[:4 * ns], dtype=np.float32).reshape(P["nsteps"], nrec, c)
    u = np.frombuffer(raw[4 * ns:], dtype=np.float64)
    # the same problem through ctypes
    M = sem.Mesh(dim, 4, physics, P["cells"], P["coords"], P["par"])
    prob = sem.Order2Newmark(M, P["dt"], P["nsteps"])
    s = P["src"]
    prob.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in P["recs"]:
        prob.add_receiver(r)
    prob.takeTimeStep(P["nsteps"])
    np.testing.assert_array_equal(seis, prob.seismograms())
    np.testing.assert_array_equal(u, prob.get_field("u").ravel())
    O = po.Sim(dim, 4, physics, P["cells"], P["coords"], P["par"])
    O.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in P["recs"]:
        O.add_receiver(r)
    ref = O.run(P["dt"], P["nsteps"])
    assert np.abs(ref).max() > 0
    assert pb.rel_l2(seis, ref) < 1e-5 and pb.rel_l2(u, O.get_field("u").ravel()) < 1e-5
