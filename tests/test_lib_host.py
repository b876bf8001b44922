"""CPU checks of the product library: it loads, exports every entry point of
include/salvus_sem.h, and its host-side gather map and colouring are
bit-exact against the oracle (no device needed)."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import meshgen, sem
from tests import refmeshes as rm

ROOT = Path(__file__).resolve().parent.parent


def test_header_and_exports_match():
    header = (ROOT / "include" / "salvus_sem.h").read_text()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sem_\w+)\s*\(", header, re.M))
    assert declared == set(sem.EXPORTS)
    L = sem.lib()
    for name in declared:
        assert hasattr(L, name), name


CASES = [
    (3, 4, lambda: meshgen.box_hex(5, 4, 3)),
    (3, 3, rm.seven_hex),
    (3, 5, rm.hex_2x2x2),
    (2, 4, lambda: meshgen.box_quad(7, 5)),
    (2, 3, rm.quad_4x4),
    (2, 9, lambda: meshgen.box_quad(3, 2)),
    (3, 1, lambda: meshgen.box_hex(2, 3, 2)),
]


@pytest.mark.parametrize("dim,order,mk", CASES)
def test_gather_map_and_colours_bit_exact(dim, order, mk):
    cells, coords = mk()
    conn, col, info = sem.topology(dim, order, cells, len(coords))
    P = po.Plex(dim, order, cells, len(coords))
    ocol, ncol = po.color_elements(dim, cells, len(coords))
    np.testing.assert_array_equal(conn, P.conn())
    np.testing.assert_array_equal(col, ocol)
    assert info["G"] == P.nglobal and info["ncolours"] == ncol
    assert (info["faces"], info["edges"]) == (P.NF, P.NEd)


def test_bad_inputs_raise():
    cells, coords = meshgen.box_hex(2, 2, 2)
    with pytest.raises(sem.SemError, match="order"):
        sem.topology(3, 9, cells, len(coords))
    bad = cells.copy()
    bad[0, 0] = 10_000
    with pytest.raises(sem.SemError, match="out of range"):
        sem.topology(3, 4, bad, len(coords))


def test_reference_step_count():
    assert sem.reference_step_count(1.0, 2e-4) == po.reference_step_count(1.0, 2e-4)
    assert sem.reference_step_count(0.2, 2e-4)[0] == 1000
