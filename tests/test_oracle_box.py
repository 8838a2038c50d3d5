"""World::run's periodic box in the oracle (src/World.cpp:1850-1901): at the start of a step a
blob whose centroid (Blob::get_centroid, src/Blob.cpp:971-983, over all N_t nodes) lies below
0 with that low wall periodic, or else above the box length with that high wall periodic, moves
by one box length (Blob::move, :935-949).  Pinned by what the mathematics fixes: the step with a
move equals the step without it translated by the box length (to rounding: the element forces
see differences of moved positions), axes and walls that do not qualify leave the blob alone, a
centroid exactly on the high wall does not move, and the moved blob's centroid is inside."""
import numpy as np
import pytest

import pyoracle
from conftest import box_case
from ffea_amd.world import linearised


def run(oracle, box, mask, shift=(0.0, 0.0, 0.0), nsteps=1):
    m, prm, sc, x0 = box_case((3, 3, 2), noise=1, sigma=0.05, seed=4, jitter=0.1)
    p = linearised(x0 * sc + np.array(shift), m.topo)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    s = oracle.stream()
    b = oracle.blob(m, prm, sc)
    b.set_solver("masslumped")
    b.set_positions(p.ravel())
    com = oracle.centroid(b)
    oracle.set_box(box(com), mask)
    try:
        for _ in range(nsteps):
            assert pyoracle.world_step(oracle, [b], s) == 0
    finally:
        oracle.set_box(np.zeros(3), 0)
    return b.get("pos").reshape(-1, 3), com, b


def test_high_and_low_walls_move_by_one_box(oracle):
    ref, com, _ = run(oracle, lambda c: np.array([1e6, 1e6, 1e6]), 0)
    scale = np.abs(ref).max()
    # x above the box: moves down by box_x; y, z inside
    bx = np.array([com[0] - 0.3, 2 * com[1], 2 * com[2]])
    got, _, b = run(oracle, lambda c: bx, 0b111111)
    assert np.allclose(got - ref, [-bx[0], 0, 0], rtol=0, atol=1e-12 * scale)
    assert 0 <= oracle.centroid(b)[0] <= bx[0]
    # the high x wall not periodic: no move at all
    got, _, _ = run(oracle, lambda c: bx, 0b111101)
    assert np.array_equal(got, ref)
    # below zero in y (blob shifted): moves up by box_y
    sh = (0.0, -com[1] - 0.2, 0.0)
    ref2, com2, _ = run(oracle, lambda c: np.array([1e6, 1e6, 1e6]), 0, shift=sh)
    by = np.array([2 * com2[0], 5.0, 2 * com2[2]])
    got, _, _ = run(oracle, lambda c: by, 0b111111, shift=sh)
    assert np.allclose(got - ref2, [0, by[1], 0], rtol=0, atol=1e-12 * scale)
    # the low y wall not periodic: nothing, even with the high one periodic
    got, _, _ = run(oracle, lambda c: by, 0b111011, shift=sh)
    assert np.array_equal(got, ref2)


def test_centroid_on_the_wall_stays(oracle):
    ref, com, _ = run(oracle, lambda c: np.array([1e6, 1e6, 1e6]), 0)
    # exactly on the high wall: not above it
    got, _, _ = run(oracle, lambda c: c.copy(), 0b111111)
    assert np.array_equal(got, ref)
    # one ulp below the centroid: above the wall, moves
    got, _, _ = run(oracle, lambda c: np.nextafter(c, -np.inf), 0b111111)
    assert np.allclose(got - ref, -np.nextafter(com, -np.inf), rtol=0, atol=1e-12 * np.abs(ref).max())


def test_centroid_is_the_sequential_mean(oracle):
    _, com, b = run(oracle, lambda c: np.array([1e6, 1e6, 1e6]), 0, nsteps=0)
    p = b.get("pos").reshape(-1, 3)
    want = np.zeros(3)
    for row in p:
        want += row
    want /= p.shape[0]
    assert np.array_equal(com, want)
