"""GPU parity of the mass-ful solvers (ffh_blob_set_solver: ConjugateGradientSolver on the
consistent mass matrix, MassLumpedSolver) against the oracle, itself bitwise equal to the
reference's own compiled solvers (tests/test_oracle_massful.py).

Bars: MassLumpedSolver in FP64 -- the whole trajectory bit-exact (element forces with the
viscous K_e u_e term, drag, noise, x = f / m, Euler update; no reduction is involved);
ConjugateGradientSolver in FP64 -- first force vector and the RngStream state bit-exact,
positions within 1e-9 of the displacement and velocities within 1e-9 relative (the CG's dot
products are summed in tree order, not in the reference's row order), same iteration counts;
mixed precision (FP32 element kernel) -- velocities within 1e-4 relative; graph replay equal to
ordinary launches bit for bit; kinetic energy within 1e-12 relative of the oracle.  The time
step is 1 fs: the explicit mass-ful update of these nanometre meshes is unstable at 10 fs, in
the reference as here (checked on the oracle).
"""
import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def case(n=(4, 3, 3), noise=1, seed=7, jit=0.1):
    from ffea_amd import units as U
    m, _, sc, x0 = box_case(n, noise=noise, sigma=0.05, seed=seed, jitter=jit)
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=noise)
    m.pinned = np.array([seed % 5, 3 * seed % 13], np.int32)
    return m, prm, sc, x0


def pair(oracle, solver, n=(4, 3, 3), noise=1, seed=7, precision="fp64"):
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case(n, noise, seed)
    sd = oracle.package_seed(seed)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    bo.set_solver(solver)
    b = Blob(m, prm, sc, rng_state=sd, precision=precision, pos=x0)
    b.set_solver(solver)
    return m, prm, sc, x0, s, bo, b


def test_lumped_trajectory_bitwise(oracle):
    import pyoracle
    m, prm, sc, x0, s, bo, b = pair(oracle, "masslumped")
    for k in range(20):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        print("lumped step", k, flush=True)
        b.step(1)
        assert np.array_equal(b.rhs().ravel(), bo.get("rhs")), k
        assert np.array_equal(b.velocities().ravel(), bo.get("vel")), k
        assert np.array_equal(b.positions().ravel(), bo.get("pos")), k
    assert np.array_equal(b.rng_state(), s.state())
    assert b.last_cg()[0] == 0
    assert b.kinetic_energy() == pytest.approx(bo.kinetic_energy(), rel=1e-12)
    assert np.abs(b.velocities()).max() > 0


def test_cg_mass_trajectory(oracle):
    import pyoracle
    m, prm, sc, x0, s, bo, b = pair(oracle, "CG")
    xs = x0.ravel() * sc
    for k in range(20):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        print("mass CG step", k, flush=True)
        b.step(1)
        if k == 0:
            assert np.array_equal(b.rhs().ravel(), bo.get("rhs"))
        po = bo.get("pos")
        assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - xs), k
        assert rel(b.velocities(), bo.get("vel")) < 1e-9, k
        assert b.last_cg()[0] == oracle.cg_iterations(bo), k
    assert np.array_equal(b.rng_state(), s.state())
    assert b.kinetic_energy() == pytest.approx(bo.kinetic_energy(), rel=1e-9)


@pytest.mark.parametrize("solver", ["CG", "masslumped"])
def test_mixed_precision(oracle, solver):
    import pyoracle
    m, prm, sc, x0, s, bo, b = pair(oracle, solver, precision="mixed")
    for k in range(5):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        b.step(1)
    assert rel(b.velocities(), bo.get("vel")) < 1e-4
    assert np.array_equal(b.rng_state(), s.state())


@pytest.mark.parametrize("solver", ["CG", "masslumped"])
def test_graph_replay_equals_ordinary_launches(solver):
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case((5, 4, 3), seed=9)
    st = np.arange(1, 7, dtype=np.uint64)
    a = Blob(m, prm, sc, rng_state=st, pos=x0)
    b = Blob(m, prm, sc, rng_state=st, pos=x0)
    a.set_solver(solver)
    b.set_solver(solver)
    a.set_graphs(False)
    for k in range(3):
        a.step(4)
        b.step(4)
        assert np.array_equal(a.positions(), b.positions()), k
        assert np.array_equal(a.velocities(), b.velocities()), k
        assert a.last_cg() == b.last_cg()


def test_batch_lumped_bitwise(oracle):
    """Two blobs in one handle, World::run's draw order (all element phases, then all position
    phases), against the oracle's world step."""
    import pyoracle
    from ffea_amd.blob import Batch
    cases = [case((3, 3, 2), seed=21), case((4, 2, 2), seed=22)]
    prm = cases[0][1]
    sd = oracle.package_seed(5)
    oracle.set_package_seed(sd)
    s = oracle.stream()
    bo = []
    for m, p, sc, x in cases:
        o = oracle.blob(m, prm, sc, pos=x)
        o.set_solver("masslumped")
        bo.append(o)
    g = Batch([c[0] for c in cases], prm, [c[2] for c in cases], rng_state=sd)
    g.set_positions(np.concatenate([c[3] * c[2] for c in cases]))
    g.set_solver("masslumped")
    for k in range(5):
        assert pyoracle.world_step(oracle, bo, s) == 0
        g.step(1)
    for got, o in zip(g.split(g.positions()), bo):
        assert np.array_equal(got.ravel(), o.get("pos"))
    assert np.array_equal(g.rng_state(), s.state())


def test_solve_given_force(oracle):
    m, prm, sc, x0, s, bo, b = pair(oracle, "CG", noise=0)
    lin = np.unique(m.topo[:, :4])
    f = np.zeros((m.num_nodes, 3))
    f[lin] = np.random.default_rng(3).normal(size=(lin.size, 3))
    st, xo = bo.solve(f.ravel())
    assert st == 0
    assert rel(b.solve(f), xo) < 1e-12
    bo.set_solver("masslumped")
    b.set_solver("masslumped")
    st, xo = bo.solve(f.ravel())
    assert np.array_equal(b.solve(f), xo)


def test_velocities_and_kinetic_energy(oracle):
    """A uniform velocity u: kinetic energy |u|^2 / 2 times the blob's mass; midpoint velocities
    come back as given."""
    m, prm, sc, x0, s, bo, b = pair(oracle, "CG", noise=0)
    u = np.array([0.3, -1.2, 0.7])
    v = np.tile(u, (m.num_nodes, 1))
    b.set_velocities(v)
    bo.set_velocities(v.ravel())
    assert np.array_equal(b.velocities(), v)
    mass = (bo.get("material")[0::5] * bo.get("vol_0")).sum()
    assert b.kinetic_energy() == pytest.approx(0.5 * mass * (u @ u), rel=1e-13)
    assert b.kinetic_energy() == pytest.approx(bo.kinetic_energy(), rel=1e-13)


def test_mass_cg_nonconvergence_is_reported():
    from ffea_amd import _lib
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case()
    prm.max_iterations_cg = 1
    prm.epsilon2 = 1e-30
    b = Blob(m, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), pos=x0)
    b.set_solver("CG")
    before = b.positions()
    with pytest.raises(_lib.CGError):
        b.step(1)
    assert np.array_equal(b.positions(), before)  # integration skipped, as the reference does


def test_mass_cg_equipartition():
    """Consistent-mass dynamics with Stokes drag and thermal noise: the mean kinetic energy of
    the second half of 4000 steps is (3/2) kT per free linear node within 15 %."""
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = case((4, 4, 3), seed=2, jit=0.0)
    m.pinned = np.zeros(0, np.int32)
    b = Blob(m, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), pos=x0)
    b.set_solver("CG")
    ke = []
    for k in range(40):
        b.step(100)
        ke.append(b.kinetic_energy())
    print("kinetic energy per 100 steps", np.round(ke, 2), flush=True)
    want = 1.5 * np.unique(m.topo[:, :4]).size * prm.kT
    assert np.mean(ke[20:]) == pytest.approx(want, rel=0.15)
