"""GPU parity of the forces from outside the element loop (ffh_blob_set_springs /
_constant_forces / _rotational_forces / _surface_forces) against the oracle (pinned in tests/test_oracle_forces.py).

Bars: with MassLumpedSolver, whose solve has no reduction, a noisy 10-step trajectory with
springs, constant forces on a vertex and on an edge midpoint, rotational forces (circular and
torque, on vertices and a midpoint, one axis through its node) and surface sets (two sharing
a face) is bit-exact -- force vector, velocities, positions, stream -- as is a two-blob batch whose
spring joins the blobs; with the massless solver the force vector is bit-exact and positions
within 1e-9 of the displacement (CG tree sums); spring energy equal to the oracle's; graph
replay equal to ordinary launches; spring ends on edge midpoints are refused.  Surface faces
include second-order sub-triangles whose corners are edge midpoints: their forces reach the
vertices through linearise_force, recomputed every step.
"""
import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu


def case(n=(4, 3, 3), seed=7):
    from ffea_amd import units as U
    m, _, sc, x0 = box_case(n, noise=1, sigma=0.05, seed=seed, jitter=0.1)
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=1)
    m.pinned = np.array([seed % 5], np.int32)
    return m, prm, sc, x0


def forces(m, x0, sc, node_off=0):
    lin = np.unique(m.topo[:, :4])
    p = x0 * sc
    n0, n1, n2 = int(lin[3]), int(lin[17]), int(lin[20])
    d = np.linalg.norm(p[n1] - p[n0])
    springs = ([n0, n2], [n1, n0], [0.8, 0.3], [0.7 * d, 1.4 * np.linalg.norm(p[n2] - p[n0])])
    mid = int(m.topo[4, 4])
    consts = ([int(lin[5]), mid, int(lin[5])], np.array([[0.3, -0.2, 0.5], [-0.4, 0.1, 0.25], [0.05, 0.0, -0.1]]))
    # vertex triangles, then second-order sub-triangles with edge midpoints as corners
    faces = np.array([m.topo[e, [0, 1, 2]] for e in range(0, 12)] + [m.topo[e, [0, 4, 5]] for e in range(12, 16)]
                     + [m.topo[e, [4, 7, 5]] for e in range(16, 18)])
    sets = [[0, 3, 5, 7], [2, 3, 11], [12, 13, 14, 15, 16, 17]]
    sf = np.array([[1.0, 0.0, 0.5], [-0.2, 0.7, 0.0], [0.0, 0.4, -0.9]])
    # rotational forces: circular and torque, on vertices (one also carrying a linear force)
    # and on an edge midpoint; one axis passes through a vertex (its own foot)
    c = p[lin].mean(axis=0)
    d = np.array([0.3, -0.4, 0.866])
    d = d / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    rmid = int(m.topo[9, 8])
    rot = ([int(lin[5]), int(lin[11]), rmid, int(lin[2])], [0.6, 0.25, 0.4, 0.3],
           np.array([list(c) + list(d), list(c) + list(d), list(c) + [1.0, 0.0, 0.0], list(p[lin[2]]) + list(d)]),
           [0, 1, 1, 0])
    return springs, (consts, rot), (faces, sets, sf)


def apply_oracle(oracle, bo, springs, consts, surf):
    oracle.set_springs([0] * len(springs[0]), springs[0], [0] * len(springs[0]), springs[1], springs[2], springs[3])
    bo.set_constant_forces(*consts[0])
    bo.set_rotational_forces(*consts[1])
    bo.set_surface_forces(*surf)


def apply_gpu(b, springs, consts, surf):
    b.set_springs(*springs)
    b.set_constant_forces(*consts[0])
    b.set_rotational_forces(*consts[1])
    b.set_surface_forces(*surf)


@pytest.fixture
def clean(oracle):
    oracle.set_springs([], [], [], [], [], [])
    yield
    oracle.set_springs([], [], [], [], [], [])


@pytest.mark.parametrize("solver", ["masslumped", "CG_nomass"])
def test_forces_trajectory(oracle, clean, solver):
    import pyoracle
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case()
    sp, ct, sf = forces(m, x0, sc)
    sd = oracle.package_seed(7)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    bo.set_solver(solver)
    apply_oracle(oracle, bo, sp, ct, sf)
    b = Blob(m, prm, sc, rng_state=sd, pos=x0)
    b.set_solver(solver)
    apply_gpu(b, sp, ct, sf)
    assert b.spring_energy() == oracle.spring_energy([bo])
    xs = x0.ravel() * sc
    for k in range(10):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        print(solver, "step", k, flush=True)
        b.step(1)
        if solver == "masslumped" or k == 0:
            assert np.array_equal(b.rhs().ravel(), bo.get("rhs")), k
        if solver == "masslumped":
            assert np.array_equal(b.positions().ravel(), bo.get("pos")), k
            assert np.array_equal(b.velocities().ravel(), bo.get("vel")), k
        else:
            po = bo.get("pos")
            assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - xs), k
    assert np.array_equal(b.rng_state(), s.state())
    assert b.spring_energy() == pytest.approx(oracle.spring_energy([bo]), rel=1e-9)


def test_batch_spring_between_blobs(oracle, clean):
    import pyoracle
    from ffea_amd.blob import Batch
    cases = [case((3, 3, 2), seed=21), case((4, 2, 2), seed=22)]
    prm = cases[0][1]
    sd = oracle.package_seed(5)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    bo = []
    for m, _, sc, x in cases:
        o = oracle.blob(m, prm, sc, pos=x)
        o.set_solver("masslumped")
        bo.append(o)
    la, lb = np.unique(cases[0][0].topo[:, :4]), np.unique(cases[1][0].topo[:, :4])
    na, nb = int(la[-1]), int(lb[0])
    # blob 1 sits on top of blob 0 in this layout; one spring joins them, one stays in blob 1
    oracle.set_springs([0, 1], [na, int(lb[2])], [1, 1], [nb, int(lb[9])], [0.5, 0.9], [1.0, 0.8])
    g = Batch([c[0] for c in cases], prm, [c[2] for c in cases], rng_state=sd)
    g.set_positions(np.concatenate([c[3] * c[2] for c in cases]))
    g.set_solver("masslumped")
    off = g.node_off
    g.set_springs([na, off[1] + int(lb[2])], [off[1] + nb, off[1] + int(lb[9])], [0.5, 0.9], [1.0, 0.8])
    assert g.spring_energy() == oracle.spring_energy(bo)
    for k in range(5):
        assert pyoracle.world_step(oracle, bo, s) == 0
        g.step(1)
    for got, o in zip(g.split(g.positions()), bo):
        assert np.array_equal(got.ravel(), o.get("pos"))
    assert np.array_equal(g.rng_state(), s.state())


def test_graph_replay_with_forces():
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case((5, 4, 3), seed=9)
    sp, ct, sf = forces(m, x0, sc)
    st = np.arange(1, 7, dtype=np.uint64)
    a = Blob(m, prm, sc, rng_state=st, pos=x0)
    b = Blob(m, prm, sc, rng_state=st, pos=x0)
    for x in (a, b):
        apply_gpu(x, sp, ct, sf)
    a.set_graphs(False)
    for k in range(3):
        a.step(4)
        b.step(4)
        assert np.array_equal(a.positions(), b.positions()), k
    # clearing every list gives back the plain step
    for x in (a, b):
        x.set_springs([], [], [], [])
        x.set_constant_forces([], np.zeros((0, 3)))
        x.set_rotational_forces([], [], np.zeros((0, 6)), [])
        x.set_surface_forces(np.zeros((0, 3)), [], np.zeros((0, 3)))
    a.step(2)
    b.step(2)
    assert np.array_equal(a.positions(), b.positions())


def test_midpoint_ends_are_refused():
    from ffea_amd import _lib
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case((3, 2, 2))
    b = Blob(m, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), pos=x0)
    lin = np.unique(m.topo[:, :4])
    mid = int(m.topo[0, 5])
    with pytest.raises(_lib.FFHError):
        b.set_springs([int(lin[0])], [mid], [1.0], [1.0])
    with pytest.raises(_lib.FFHError):
        b.set_springs([int(lin[0])], [int(lin[0])], [1.0], [1.0])
    ax = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 1.0]])
    with pytest.raises(_lib.FFHError):  # neither 'f' (0) nor 't' (1)
        b.set_rotational_forces([int(lin[0])], [1.0], ax, [2])
    with pytest.raises(_lib.FFHError):
        b.set_rotational_forces([m.num_nodes], [1.0], ax, [0])
