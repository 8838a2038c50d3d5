"""The HIP path on the benchmark workloads at their full sizes (bench.py's make_case), against the
oracle where its cost allows and through size-independent properties where it does not.

c1, the 48^3 Kuhn cube (117,649 linear nodes, 663,552 tets, noise on, seed 44):
  - first step: the force vector handed to the solver (element forces + every RngStream draw)
    and the stream state after it bit-exact against the oracle, the solution within 1e-9;
  - 10 noise-off steps: FP64 positions within 1e-10 of the oracle's, relative to the
    displacement (the CG dot products are summed in tree order, the reference's sequentially);
  - the mixed path over the same 10 steps: positions within 1e-5 of the displacement,
    velocities within 1e-5 of the FP64 path's.
c2, 256 uneven jittered blobs with scaled materials (10.9 M tets), one batch:
  - three blobs' force vectors bit-exact against the oracle, each blob's element noise drawn after
    the earlier blobs' element draws and its node noise after every blob's element draws and the
    earlier blobs' node draws (World::run's order, src/World.cpp:2020-2041), on both paths
    (the mixed one within 1e-5); every blob converges on every step;
  - step graphs replayed against ordinary launches from the same state: bit-identical.
c4, the 192x96x48 jittered box with a 10x stiffer core (917,329 linear nodes, 5.3 M tets):
  - every step converges on both paths, the mixed one within 1e-4 of the FP64 velocities;
  - the assembled matrix on a uniform translation is the Stokes drag times it (the viscosity
    matrix annihilates rigid translation, src/tetra_element_linear.cpp:299-328);
  - a 24x12x6 box from the same generator (jitter, stiff core) against the oracle over 3 noisy
    steps: the first force vector bit-exact, later ones within 1e-9, positions within 1e-9 of
    the displacement, the same CG iteration counts and the stream bit-exact.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(cfg, **prm_over):
    import bench
    ms, prm, sc, xs, _ = bench.make_case(cfg)
    for k, v in prm_over.items():
        setattr(prm, k, v)
    return ms, prm, sc, xs


# ---------------------------------------------------------------------------------------- c1
def test_c1_first_step_bitwise(oracle):
    import pyoracle
    from ffea_amd.blob import Blob
    ms, prm, sc, xs = make("c1")
    m, x0 = ms[0], xs[0]
    seed = oracle.package_seed(prm.rng_seed)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    print("c1 oracle blob", m.num_elements, "tets", flush=True)
    bo = oracle.blob(m, prm, sc, pos=x0)
    assert pyoracle.world_step(oracle, [bo], s) == 0
    b = Blob(m, prm, sc, rng_state=seed, precision="fp64", pos=x0)
    b.step(1)
    assert np.array_equal(b.rhs().ravel(), bo.get("rhs")), "force vector not bit-exact"
    assert np.array_equal(b.rng_state(), s.state())
    assert rel(b.velocities(), bo.get("vel")) < 1e-9
    po = bo.get("pos")
    assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - x0.ravel() * sc)
    assert b.last_cg()[0] == oracle.cg_iterations(bo)


def test_c1_noise_off_trajectory(oracle):
    import pyoracle
    from ffea_amd.blob import Blob
    ms, prm, sc, xs = make("c1", calc_noise=0)
    m, x0 = ms[0], xs[0]
    seed = oracle.package_seed(prm.rng_seed)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    a = Blob(m, prm, sc, rng_state=seed, precision="fp64", pos=x0)
    b = Blob(m, prm, sc, rng_state=seed, precision="mixed", pos=x0)
    xs0 = x0.ravel() * sc
    for k in range(10):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        a.step(1)
        b.step(1)
        po = bo.get("pos")
        d = np.linalg.norm(po - xs0)
        assert np.linalg.norm(a.positions().ravel() - po) <= 1e-10 * d, k
        assert np.linalg.norm(b.positions().ravel() - po) <= 1e-5 * d, k
        assert rel(b.velocities(), a.velocities()) < 1e-5, k
        print("c1 step", k, "fp64", np.linalg.norm(a.positions().ravel() - po) / d,
              "mixed", np.linalg.norm(b.positions().ravel() - po) / d, flush=True)


# ---------------------------------------------------------------------------------------- c2
@pytest.fixture(scope="module")
def c2():
    return make("c2")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c2_batch_blob_forces_and_convergence(oracle, c2, precision):
    from ffea_amd.blob import Batch
    ms, prm, sc, xs = c2
    seed = oracle.package_seed(prm.rng_seed)
    g = Batch(ms, prm, sc, rng_state=seed, precision=precision)
    g.set_positions(np.concatenate(xs) * sc)
    print("c2", precision, g.sizes(), flush=True)
    g.update_internal_forces()
    g.update_positions()
    its, res, stat = g.last_cg_per_blob()
    assert (stat == 0).all() and (res < prm.epsilon2).all()
    rhs = g.split(g.rhs())
    e_tot = sum(m.num_elements for m in ms)
    for k in (0, 127, 255):
        elem_before = sum(7 * m.num_elements for m in ms[:k])
        node_before = 7 * e_tot + sum(3 * m.num_nodes for m in ms[:k])
        oracle.set_package_seed(seed)
        se = oracle.stream()
        se.advance(0, elem_before)
        oracle.set_package_seed(seed)
        sn = oracle.stream()
        sn.advance(0, node_before)
        bo = oracle.blob(ms[k], prm, sc, pos=xs[k])
        inv, want = bo.forces_split(se, sn)
        assert inv == 0
        if precision == "fp64":
            assert np.array_equal(rhs[k].ravel(), want), f"blob {k}"
        else:
            assert rel(rhs[k], want) < 1e-5, f"blob {k}"
        bo.close()
    for step in range(2):
        g.step(1)
        its, res, stat = g.last_cg_per_blob()
        assert (stat == 0).all() and (res < prm.epsilon2).all(), step
    print("c2", precision, "iterations per blob: mean", its.mean(), "max", its.max(), flush=True)


def test_c2_graph_replay_equals_ordinary_launches(c2):
    from ffea_amd.blob import Batch, package_seed
    ms, prm, sc, xs = c2
    g = Batch(ms, prm, sc, rng_state=package_seed(prm.rng_seed), precision="mixed")
    x0 = np.concatenate(xs) * sc
    g.set_positions(x0)
    st0 = g.rng_state()
    g.step(3)
    pa, sa = g.positions(), g.rng_state()
    g.set_positions(x0)
    g.set_rng_state(st0)
    g.set_graphs(False)
    g.step(3)
    assert np.array_equal(g.positions(), pa)
    assert np.array_equal(g.rng_state(), sa)


# ---------------------------------------------------------------------------------------- c4
def test_c4_full_size_properties():
    from ffea_amd.blob import Blob
    ms, prm, sc, xs = make("c4")
    m, x0 = ms[0], xs[0]
    seed = np.arange(1, 7, dtype=np.uint64)
    a = Blob(m, prm, sc, rng_state=seed, precision="fp64", pos=x0)
    print("c4", a.sizes(), flush=True)
    # K u = zeta u for a uniform translation u of the linear nodes
    assert a.update_internal_forces() == 0
    u = np.zeros((m.num_nodes, 3))
    u[: m.num_linear] = [0.3, -1.1, 0.7]
    Ku = a.apply_matrix(u).reshape(-1, 3)[: m.num_linear]
    zeta = 6 * np.pi * prm.stokes_visc * m.stokes_radius[: m.num_linear] * sc
    want = zeta[:, None] * u[: m.num_linear]
    assert np.max(np.abs(Ku - want)) <= 1e-10 * np.max(np.abs(want))
    b = Blob(m, prm, sc, rng_state=seed, precision="mixed", pos=x0)
    for k in range(3):  # noisy steps from one state: the same draws, solves of nearby matrices
        a.step(1)
        b.step(1)
        assert rel(b.velocities(), a.velocities()) < 1e-4, k
        print("c4 step", k, "iterations", a.last_cg()[0], b.last_cg()[0], flush=True)


def test_c4_generator_small_box_against_oracle(oracle):
    import pyoracle
    from ffea_amd import mesh as M, units as U
    from ffea_amd.blob import Blob
    m = M.box_mesh(24, 12, 6, jitter=0.15, seed=4)
    M.stiff_core(m, axis=0, factor=10.0)
    x0 = M.perturb(m, 0.02, seed=1)
    prm = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=1, calc_stokes=1, rng_seed=44)
    sc = U.blob_scale(1e-9)
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    b = Blob(m, prm, sc, rng_state=seed, precision="fp64", pos=x0)
    xs0 = x0.ravel() * sc
    for k in range(3):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        b.step(1)
        if k == 0:  # later steps start from positions equal to rounding (tree-order CG sums)
            assert np.array_equal(b.rhs().ravel(), bo.get("rhs"))
        else:
            assert rel(b.rhs(), bo.get("rhs")) < 1e-9, k
        po = bo.get("pos")
        assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - xs0), k
        assert b.last_cg()[0] == oracle.cg_iterations(bo), k
    assert np.array_equal(b.rng_state(), s.state())
