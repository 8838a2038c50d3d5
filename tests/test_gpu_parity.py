"""GPU parity: the HIP path (through the C-ABI) against the reference's golden vectors and the
oracle.

Bars: FP64 path -- element forces, Stokes noise, gathered force vector and RngStream state
bit-exact; velocities (CG solutions, whose dot products are summed in tree order instead of the
reference's sequential order) within 1e-9 relative in norm; positions within 1e-12 relative.
Mixed path (FP32 element/assembly/SpMV storage, FP64 CG scalars, RHS and integration) --
force vector within 1e-5 relative in norm, velocities within 1e-4.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, box_case
from test_golden import CASES, load_case

pytestmark = pytest.mark.gpu


def gpu_blob(m, prm, sc, seed, x0, precision="fp64"):
    from ffea_amd.blob import Blob
    return Blob(m, prm, sc, rng_state=seed, precision=precision, pos=x0)


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def seq_csr_apply(key, col, val, x):
    """SparseMatrixFixedPattern::apply's sequential row sums, in plain Python doubles."""
    out = np.zeros(key.size - 1)
    xl = x.tolist()
    vl, cl = val.tolist(), col.tolist()
    for i in range(key.size - 1):
        w = 0.0
        for j in range(key[i], key[i + 1]):
            w += vl[j] * xl[cl[j]]
        out[i] = w
    return out


@pytest.mark.parametrize("name", CASES)
def test_fp64_first_step_bitwise(name):
    g, m, prm, sc = load_case(name)
    print(f"case {name}", flush=True)
    b = gpu_blob(m, prm, sc, g["seed"], g["x0"])
    assert b.update_internal_forces() == 0
    b.update_positions()
    assert np.array_equal(b.rhs().ravel(), g["rhs"]), "force vector (element forces + noise) not bit-exact"
    # RNG consumed exactly one step's draws
    assert np.array_equal(b.rng_state(), g["rng_after_forces"])
    assert rel(b.positions(), g["traj"][0]) < 1e-12


@pytest.mark.parametrize("name", CASES)
def test_fp64_three_steps(name):
    g, m, prm, sc = load_case(name)
    b = gpu_blob(m, prm, sc, g["seed"], g["x0"])
    x0 = g["x0"].ravel() * sc
    for k in range(g["traj"].shape[0]):
        b.step(1)
        p = b.positions().ravel()
        # error relative to the displacement accumulated so far, and to |x|
        assert np.linalg.norm(p - g["traj"][k]) <= 1e-9 * np.linalg.norm(g["traj"][k] - x0) + 1e-15 * np.linalg.norm(x0)
    assert rel(b.velocities(), g["vel"]) < 1e-9
    assert np.array_equal(b.rng_state(), g["rng_after_steps"])


@pytest.mark.parametrize("name", CASES)
def test_fp64_matrix_and_solve(name):
    g, m, prm, sc = load_case(name)
    b = gpu_blob(m, prm, sc, g["seed"], g["x0"])
    b.update_internal_forces()  # element matrices at x0, as in the golden
    x = np.random.default_rng(3).normal(size=3 * m.num_nodes)
    want = seq_csr_apply(g["csr_key"], g["csr_col"], g["csr_val"], x)
    got = b.apply_matrix(x)
    assert np.array_equal(got, want), "assembled matrix or SpMV order differs from the reference"
    v = b.solve(g["rhs"])
    assert rel(v, g["solve_v"]) < 1e-9


@pytest.mark.parametrize("name", CASES)
def test_mixed_precision_against_golden(name):
    g, m, prm, sc = load_case(name)
    b = gpu_blob(m, prm, sc, g["seed"], g["x0"], precision="mixed")
    b.update_internal_forces()
    b.update_positions()
    assert rel(b.rhs(), g["rhs"]) < 1e-5
    assert np.array_equal(b.rng_state(), g["rng_after_forces"])
    b.step(2)
    assert rel(b.velocities(), g["vel"]) < 1e-4
    x0 = g["x0"].ravel() * sc
    assert np.linalg.norm(b.positions().ravel() - g["traj"][2]) <= 1e-4 * np.linalg.norm(g["traj"][2] - x0)


def test_fp64_c0_trajectory_against_oracle(oracle):
    """C0 cube (8^3 cells, 729 linear nodes, 3072 tets), noise on, 20 steps."""
    import pyoracle
    m, prm, sc, x0 = box_case((8, 8, 8), noise=1)
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    b = gpu_blob(m, prm, sc, seed, x0)
    xs = x0.ravel() * sc
    for k in range(20):
        assert pyoracle.world_step(oracle, [bo], s) == 0
        b.step(1)
        if k == 0:
            assert np.array_equal(b.rhs().ravel(), bo.get("rhs"))
        po = bo.get("pos")
        assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - xs)
        assert b.last_cg()[0] == oracle.cg_iterations(bo)
    assert np.array_equal(b.rng_state(), s.state())


def test_fp64_noise_off_long_trajectory(oracle):
    """Noise-off relaxation, 1000 steps: FP64 GPU within 1e-10 relative of the oracle."""
    import pyoracle
    m, prm, sc, x0 = box_case((4, 4, 3), noise=0, sigma=0.05, seed=2)
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    b = gpu_blob(m, prm, sc, seed, x0)
    for _ in range(1000):
        assert pyoracle.world_step(oracle, [bo], s) == 0
    b.step(1000)
    po = bo.get("pos")
    xs = x0.ravel() * sc
    assert np.linalg.norm(b.positions().ravel() - po) <= 1e-10 * np.linalg.norm(po - xs)
    assert np.linalg.norm(b.positions().ravel() - po) <= 1e-10 * np.linalg.norm(po)


def test_inversion_is_reported():
    """Moving one node through the opposite face flips an element (src/tetra_element_linear.cpp:260)."""
    from ffea_amd import _lib
    m, prm, sc, x0 = box_case((2, 2, 2), noise=0)
    b = gpu_blob(m, prm, sc, np.array([1, 2, 3, 4, 5, 6], np.uint64), x0)
    p = b.positions()
    p[0] += 3.0 * 5.9  # node 0 pushed through its elements
    b.set_positions(p.ravel())
    with pytest.raises(_lib.InversionError):
        b.step(1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_nonconvergence_is_reported(precision):
    from ffea_amd import _lib
    m, prm, sc, x0 = box_case((3, 3, 3), noise=1)
    prm.max_iterations_cg = 1
    prm.epsilon2 = 1e-30
    b = gpu_blob(m, prm, sc, np.array([1, 2, 3, 4, 5, 6], np.uint64), x0, precision)
    before = b.positions()
    with pytest.raises(_lib.CGError):
        b.step(1)
    assert np.array_equal(b.positions(), before)  # integration skipped, as the reference does


def test_strain_energy_against_oracle(oracle):
    """k_strain_energy (Blob::make_measurements' sum, src/Blob.cpp:1424-1432) against the oracle
    on a jittered, displaced box; the sum order differs (block tree), hence the tolerance."""
    m, prm, sc, x0 = box_case((4, 3, 3), noise=0, sigma=0.05, seed=6, jitter=0.1)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    assert bo.forces_only(s)[0] == 0
    b = gpu_blob(m, prm, sc, np.arange(1, 7, dtype=np.uint64), x0)
    assert b.update_internal_forces() == 0
    assert b.strain_energy() == pytest.approx(oracle.strain_energy(bo), rel=1e-12)


def test_large_blob_per_blob_totals(oracle):
    """A blob of more than 3072 slices (61 x 59 x 59 nodes) takes the per-blob totals path of the
    CG reductions: one noisy step against the oracle, same bars as the c0 trajectory."""
    import pyoracle
    m, prm, sc, x0 = box_case((60, 58, 58), noise=1)
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    b = gpu_blob(m, prm, sc, seed, x0)
    assert b.sizes()["num_linear"] > 3072 * 64
    print("large blob", b.sizes(), flush=True)
    assert pyoracle.world_step(oracle, [bo], s) == 0
    b.step(1)
    assert np.array_equal(b.rhs().ravel(), bo.get("rhs"))
    po = bo.get("pos")
    assert np.linalg.norm(b.positions().ravel() - po) <= 1e-9 * np.linalg.norm(po - x0.ravel() * sc)
    assert b.last_cg()[0] == oracle.cg_iterations(bo)
    assert np.array_equal(b.rng_state(), s.state())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_ordinary_launches(precision):
    """ffh_blob_step through a replayed step graph against the same steps launched one by one:
    bit-identical positions and RngStream state after every step, with host reads and
    synchronisations between the replays; then a stretch of steps in one call."""
    m, prm, sc, x0 = box_case((5, 4, 3), noise=1, jitter=0.1)
    seed = np.arange(3, 9, dtype=np.uint64)
    a = gpu_blob(m, prm, sc, seed, x0, precision)
    b = gpu_blob(m, prm, sc, seed, x0, precision)
    b.set_graphs(False)
    for k in range(6):
        a.step(1)
        b.step(1)
        assert np.array_equal(a.positions(), b.positions()), k
        assert np.array_equal(a.rng_state(), b.rng_state()), k
        assert a.last_cg() == b.last_cg()
    a.step(25)
    b.step(25)
    assert np.array_equal(a.positions(), b.positions())
    assert a.steps_done() == b.steps_done() == 31


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("batch", [False, True])
def test_tuned_launch_shapes_give_identical_results(precision, batch):
    """Every launch shape the tuner may pick (element / assembly workgroup size, assembly unroll
    depth, SpMV waves per slice) gives bit-identical trajectories, for one blob and for a batch."""
    from ffea_amd.blob import Batch
    seed = np.arange(5, 11, dtype=np.uint64)
    if batch:
        from test_gpu_batch import uneven_blobs
        meshes, x0s, prm, sc = uneven_blobs()

        def make():
            g = Batch(meshes, prm, sc, rng_state=seed, precision=precision)
            g.set_positions(np.concatenate(x0s) * sc)
            return g
    else:
        m, prm, sc, x0 = box_case((6, 5, 4), noise=1, jitter=0.1)

        def make():
            return gpu_blob(m, prm, sc, seed, x0, precision)
    ref = None
    for shape in ((256, 512, 2, 3, 0), (64, 256, 1, 1, 1), (128, 512, 4, 3, 1), (256, 256, 4, 1, 0),
                  (128, 512, 1, 1, 1), (64, 256, 2, 3, 0), (64, 1024, 1, 3, 0), (128, 1024, 2, 1, 1)):
        b = make()
        b.set_tuning(*shape)
        assert b.tuning() == dict(zip(b.TUNING_KEYS, shape))
        b.step(6)
        out = (b.positions(), b.rng_state(), b.last_cg())
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]) and out[2] == ref[2], shape


def test_uniform_external_force_moves_blob_at_stokes_velocity():
    """A uniform external force F on every vertex of a free blob at rest, noise off: the
    viscosity matrix annihilates rigid translation and the rest state has no elastic force, so
    every vertex moves at v = F / zeta (Stokes' law, zeta = 6 pi eta r), as a reference run
    with the same face forces would."""
    m, prm, sc, x0 = box_case((4, 3, 3), noise=0)
    b = gpu_blob(m, prm, sc, np.arange(1, 7, dtype=np.uint64), m.pos)
    nl = m.num_linear
    F = np.zeros((m.num_nodes, 3))
    F[:nl] = [1e-3, -2e-3, 5e-4]
    b.set_external_forces(F)
    b.step(1)
    zeta = 6 * np.pi * prm.stokes_visc * m.stokes_radius[:nl] * sc
    v = b.velocities()[:nl]
    assert rel(v, F[:nl] / zeta[:, None]) < 1e-5
    p = b.positions()[:nl]
    # a displacement of ~1e-7 read off coordinates of ~20: ~1e-8 relative is the rounding floor
    assert rel(p - m.pos[:nl] * sc, v * prm.dt) < 1e-7
    b.set_external_forces(None)
    b.step(1)
    assert np.abs(b.velocities()[:nl]).max() < 1e-3 * np.abs(v).max()


def test_midpoint_external_force_goes_to_edge_ends():
    """A force on one edge midpoint reaches the solver as half of it on each end of the edge
    (linearise_force, src/Blob.cpp:1066-1085), nowhere else."""
    from ffea_amd.mesh import EDGES
    m, prm, sc, x0 = box_case((3, 3, 2), noise=0)
    seed = np.arange(1, 7, dtype=np.uint64)
    a = gpu_blob(m, prm, sc, seed, x0)
    b = gpu_blob(m, prm, sc, seed, x0)
    e, slot = 7, 3  # element 7, its edge (1, 2)
    mid = m.topo[e, 4 + slot]
    ends = [m.topo[e, EDGES[slot][0]], m.topo[e, EDGES[slot][1]]]
    F = np.zeros((m.num_nodes, 3))
    F[mid] = [2e-3, -1e-3, 4e-3]
    b.set_external_forces(F)
    a.step(1)
    b.step(1)
    d = b.rhs() - a.rhs()
    scale = np.abs(a.rhs()).max()
    for n in ends:
        assert np.allclose(d[n], 0.5 * F[mid], rtol=0, atol=1e-12 * scale)
    others = np.setdiff1d(np.arange(m.num_nodes), ends)
    assert np.abs(d[others]).max() <= 1e-12 * scale


@pytest.mark.parametrize("period", [0, 3, 8])
def test_mixed_cg_stops_on_the_true_residual(period, monkeypatch):
    """The mixed path's CG (FP32 matrix and work vectors; FP64 iterate, dot products and true
    residual) stops only when |f - K x|^2 / |f|^2 < eps^2 for the FP64 residual of its own FP32
    matrix, formed every `period` iterations (0: only when the FP32 recurrence passes) and at
    every candidate stop.  K x is checked independently through apply_matrix (the same FP32
    blocks, rows summed in FP64).  Its solution agrees with the FP64 path's within 2e-5: both
    meet eps = 1e-6 on matrices that differ by FP32 rounding, with kappa(D^-1 K) of a few."""
    monkeypatch.setenv("FFH_RR_PERIOD", str(period))
    m, prm, sc, x0 = box_case((6, 5, 4), noise=1, jitter=0.1, seed=5)
    seed = np.arange(1, 7, dtype=np.uint64)
    f = np.random.default_rng(7).normal(size=(m.num_nodes, 3))
    f[m.num_linear:] = 0.0
    f = f.ravel()
    out = {}
    for prec in ("fp64", "mixed"):
        b = gpu_blob(m, prm, sc, seed, x0, prec)
        assert b.update_internal_forces() == 0
        v = b.solve(f)
        it, res = b.last_cg()
        Kv = b.apply_matrix(v)
        true = np.sum((f - Kv) ** 2) / np.sum(f ** 2)
        print(prec, "period", period, "iterations", it, "reported", res, "true", true, flush=True)
        assert true < prm.epsilon2
        if prec == "mixed":  # the stopping test was on the true residual itself
            assert res == pytest.approx(true, rel=1e-9)
        out[prec] = v
    assert rel(out["mixed"], out["fp64"]) < 2e-5


def test_mixed_cg_trajectory_against_fp64():
    """Noisy steps on the mixed path against the FP64 path from the same state: the force
    vectors agree to FP32 element rounding, velocities within 1e-4, and every solve converges."""
    m, prm, sc, x0 = box_case((8, 8, 8), noise=1)
    seed = np.arange(2, 8, dtype=np.uint64)
    a = gpu_blob(m, prm, sc, seed, x0, "fp64")
    b = gpu_blob(m, prm, sc, seed, x0, "mixed")
    xs = x0.ravel() * sc
    for k in range(10):
        a.step(1)
        b.step(1)
        assert rel(b.rhs(), a.rhs()) < 1e-5, k
        assert rel(b.velocities(), a.velocities()) < 1e-4, k
        pa = a.positions().ravel()
        assert np.linalg.norm(b.positions().ravel() - pa) <= 1e-4 * np.linalg.norm(pa - xs), k
    assert np.array_equal(a.rng_state(), b.rng_state())


@pytest.mark.parametrize("noise", [0, 1])
def test_stress_magnitudes_against_oracle(oracle, noise):
    """internal_stress_mag per element (src/Blob.cpp:700) from the element kernel: FP64 bit-exact
    against the oracle (itself bit-exact against the reference's objects), with and without the
    fluctuating stress; the mixed path within FP32 rounding."""
    m, prm, sc, x0 = box_case((4, 3, 3), noise=noise, sigma=0.05, seed=8, jitter=0.1)
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    assert bo.forces_only(s)[0] == 0
    want = bo.get("stress_mag")
    assert want.min() > 0
    for prec in ("fp64", "mixed"):
        b = gpu_blob(m, prm, sc, seed, x0, prec)
        b.set_stress_output(True)
        assert b.update_internal_forces() == 0
        got = b.stress_magnitudes()
        if prec == "fp64":
            assert np.array_equal(got, want)
        else:
            assert rel(got, want) < 1e-5
        b.set_stress_output(False)
        b.step(1)  # steps without the output keep working (graphs rebuilt without it)
