"""ffh_blob_ssint_energy_noalloc (the interaction energy with no device allocation: pair
energies in the library's per-GPU scratch array) against ffh_blob_ssint_energy on the same
handle, through the C-ABI, bit for bit: a larger workload, then a smaller one right after it
(the scratch keeps the first call's contents), and a workload with more face pairs than the
scratch holds, which must be refused while the allocating entry point still serves it.

The workloads are the reference's two-cube fixture (tests/physics/cubes_lj) with the
Lennard-Jones interaction: every face of both cubes active with self-interaction on a grid of
80 cells of h/8 per axis (at most about 4 * 10^5 pairs); the cubes_lj set-up (the facing top
faces only, at most 4 * 10^4 pairs); every face active with self-interaction on two cells of
44 per axis (neighbour offsets wrapping onto the same cells: 7,332,409 pairs).
"""
import ctypes as C

import numpy as np
import pytest

from test_gpu_ssint import cubes

pytestmark = pytest.mark.gpu


def _energies(g):
    from ffea_amd import _lib
    a, b = C.c_double(), C.c_double()
    ra = g._L.ffh_blob_ssint_energy(g._h, C.byref(a))
    rb = g._L.ffh_blob_ssint_energy_noalloc(g._h, C.byref(b))
    return ra, a.value, rb, b.value, _lib.last_error()


def _workload(oracle, all_faces, grid, cell, inc):
    """cell: the grid's cell size in meso units, or a divisor of the 3 nm cutoff (int)."""
    g, _, _, faces, types, E, R, h, opp = cubes(oracle, all_faces=all_faces)
    g.set_ssint(faces, types, E, R, grid, h / cell if isinstance(cell, int) else cell, inc, 1)
    g.update_internal_forces()  # the solve whose pairs the energy sums
    return g, int((types >= 0).sum())


def test_noalloc_energy_equals_allocating_entry_point(oracle):
    big, nb = _workload(oracle, True, (80, 80, 80), 8, 1)
    small, ns = _workload(oracle, False, (10, 10, 10), 1, 0)
    assert ns < nb
    for g, name in ((big, "all faces, self, 80 cells"), (small, "cubes_lj top faces")):
        print(name, flush=True)
        ra, a, rb, b, msg = _energies(g)
        assert ra == 0 and rb == 0, msg
        assert a != 0.0
        assert a == b, (name, a, b)
    # the big workload again after the small one: nothing of the small call is read back
    ra, a, rb, b, msg = _energies(big)
    assert ra == 0 and rb == 0, msg
    assert a == b


def test_noalloc_energy_refuses_more_pairs_than_its_scratch(oracle):
    g, _ = _workload(oracle, True, (2, 2, 2), 44.0, 1)
    print("two cells per axis, every face, self-interaction", flush=True)
    ra, a, rb, b, msg = _energies(g)
    assert ra == 0 and a != 0.0
    from ffea_amd import _lib
    assert rb == _lib.FFH_E_ARG
    assert "no-allocation scratch" in msg
