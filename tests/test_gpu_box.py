"""GPU parity of World::run's periodic box (ffh_blob_set_box) against the oracle (pinned in
tests/test_oracle_box.py).  With MassLumpedSolver (no reductions in the solve) noisy
trajectories are bit-exact -- positions and stream every step -- for a blob leaving through a
high wall, one below a low wall, a centroid one ulp past the wall and one exactly on it (both
decided by the sequential file-order sum), walls that are not periodic, and a two-blob batch
where only one blob moves.  Graph replay equals ordinary launches; removing the box restores
the plain step."""
import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu


def setup(oracle, n=(4, 3, 3), seed=7, shift=(0.0, 0.0, 0.0)):
    import pyoracle  # noqa: F401
    from ffea_amd import units as U
    from ffea_amd.blob import Blob
    from ffea_amd.world import linearised
    m, _, sc, x0 = box_case(n, noise=1, sigma=0.05, seed=seed, jitter=0.1)
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=1)
    p = linearised(x0 * sc + np.array(shift), m.topo)
    sd = oracle.package_seed(seed)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc)
    bo.set_solver("masslumped")
    bo.set_positions(p.ravel())
    b = Blob(m, prm, sc, rng_state=sd)
    b.set_positions(p)
    b.set_solver("masslumped")
    return b, bo, s


def compare(oracle, b, bo, s, box, mask, nsteps=6):
    import pyoracle
    oracle.set_box(box, mask)
    b.set_box(box, mask)
    try:
        for k in range(nsteps):
            assert pyoracle.world_step(oracle, [bo], s) == 0
            b.step(1)
            print("step", k, flush=True)
            assert np.array_equal(b.positions().ravel(), bo.get("pos")), k
    finally:
        oracle.set_box(np.zeros(3), 0)
    assert np.array_equal(b.rng_state(), s.state())


@pytest.mark.parametrize("case", ["high_x", "low_y", "ulp_past", "on_wall", "walls_off"])
def test_wrap_trajectory(oracle, case):
    shift = (0.0, -30.0, 0.0) if case == "low_y" else (0.0, 0.0, 0.0)
    b, bo, s = setup(oracle, shift=shift)
    com = oracle.centroid(bo)
    big = 4 * np.abs(com) + 10
    mask = 0b111111
    if case == "high_x":
        box = np.array([com[0] - 0.3, big[1], big[2]])
    elif case == "low_y":
        box = np.array([big[0], 5.0, big[2]])
    elif case == "ulp_past":
        box = np.array([big[0], np.nextafter(com[1], -np.inf), big[2]])
    elif case == "on_wall":
        box = np.array([big[0], big[1], com[2]])
    else:
        box = np.array([com[0] - 0.3, big[1], big[2]])
        mask = 0b111101
    print(case, com, box, flush=True)
    compare(oracle, b, bo, s, box, mask)


def test_batch_moves_one_blob(oracle):
    import pyoracle
    from ffea_amd import units as U
    from ffea_amd.blob import Batch
    from ffea_amd.world import linearised
    cases = [box_case((3, 3, 2), seed=21), box_case((4, 2, 2), seed=22)]
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=1)
    sd = oracle.package_seed(5)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    ps = [linearised(x * sc + np.array([0.0, 0.0, 40.0 * k]), m.topo) for k, (m, _, sc, x) in enumerate(cases)]
    bo = []
    for (m, _, sc, _), p in zip(cases, ps):
        o = oracle.blob(m, prm, sc)
        o.set_solver("masslumped")
        o.set_positions(p.ravel())
        bo.append(o)
    g = Batch([c[0] for c in cases], prm, [c[2] for c in cases], rng_state=sd)
    g.set_positions(np.concatenate(ps))
    g.set_solver("masslumped")
    coms = [oracle.centroid(o) for o in bo]
    box = np.array([60.0, 60.0, 0.5 * (coms[0][2] + coms[1][2])])  # the second blob is above it
    oracle.set_box(box, 0b111111)
    g.set_box(box)
    try:
        for k in range(4):
            assert pyoracle.world_step(oracle, bo, s) == 0
            g.step(1)
            for got, o in zip(g.split(g.positions()), bo):
                assert np.array_equal(got.ravel(), o.get("pos")), k
    finally:
        oracle.set_box(np.zeros(3), 0)
    assert g.split(g.positions())[1][:, 2].mean() < box[2]


def test_graph_replay_and_removal(oracle):
    b, _, _ = setup(oracle, seed=9)
    a, _, _ = setup(oracle, seed=9)
    com = np.asarray(b.positions()).mean(axis=0)
    box = np.array([com[0] - 0.3, 4 * abs(com[1]) + 10, 4 * abs(com[2]) + 10])
    for x in (a, b):
        x.set_box(box)
    a.set_graphs(False)
    for k in range(3):
        a.step(3)
        b.step(3)
        assert np.array_equal(a.positions(), b.positions()), k
    for x in (a, b):
        x.set_box(None)
    a.step(2)
    b.step(2)
    assert np.array_equal(a.positions(), b.positions())
