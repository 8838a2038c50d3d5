"""GPU parity of a batch: several blobs of uneven size stepped together (ffh_batch_create) on one
RngStream, against the oracle's World::run over the same blobs (src/World.cpp:2020-2041: every
blob's element phase, then every blob's position phase, one stream).

Bars as for one blob: the force vector of every blob and the stream state bit-exact, positions
within 1e-9 of the displacement, the same CG iteration count per blob.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def uneven_blobs(noise=1, eps=1e-6):
    """Three boxes of different shape: one jittered with a pinned face, one with rescaled
    material, one plain; seeded displacements."""
    from ffea_amd import mesh as M, units as U
    meshes, x0s = [], []
    for k, (n, jit) in enumerate((((3, 3, 2), 0.0), ((4, 2, 2), 0.15), ((2, 3, 5), 0.1))):
        m = M.box_mesh(*n, jitter=jit, seed=10 + k)
        if k == 1:
            lin = m.pos[: m.num_linear]
            m.pinned = np.nonzero(lin[:, 0] < lin[:, 0].min() + 0.5)[0].astype(np.int32)
        if k == 2:
            M.scale_material(m, 0.5, 2.0, seed=7)
        meshes.append(m)
        x0s.append(M.perturb(m, 0.02, seed=20 + k))
    prm = U.SimParams.from_si(epsilon=eps, calc_noise=noise)
    return meshes, x0s, prm, U.blob_scale(1e-9)


def make_pair(oracle, meshes, x0s, prm, sc, precision="fp64"):
    from ffea_amd.blob import Batch
    seed = oracle.package_seed(44)
    oracle.set_package_seed(seed)
    s = oracle.stream()
    bos = [oracle.blob(m, prm, sc, pos=x) for m, x in zip(meshes, x0s)]
    g = Batch(meshes, prm, sc, rng_state=seed, precision=precision)
    g.set_positions(np.concatenate(x0s) * sc)
    return s, bos, g


def test_batch_fp64_against_oracle(oracle):
    import pyoracle
    meshes, x0s, prm, sc = uneven_blobs()
    s, bos, g = make_pair(oracle, meshes, x0s, prm, sc)
    print("batch of", [m.num_nodes for m in meshes], "nodes", flush=True)
    xs = [x.ravel() * sc for x in x0s]
    for k in range(10):
        assert pyoracle.world_step(oracle, bos, s) == 0
        g.step(1)
        its, res, stat = g.last_cg_per_blob()
        assert (stat == 0).all()
        if k == 0:
            for f, bo in zip(g.split(g.rhs()), bos):
                assert np.array_equal(f.ravel(), bo.get("rhs")), "force vector not bit-exact"
        for i, (p, bo) in enumerate(zip(g.split(g.positions()), bos)):
            po = bo.get("pos")
            assert np.linalg.norm(p.ravel() - po) <= 1e-9 * np.linalg.norm(po - xs[i]), f"blob {i}"
            assert its[i] == oracle.cg_iterations(bo), f"blob {i}"
    assert np.array_equal(g.rng_state(), s.state())


def test_batch_mixed_against_oracle(oracle):
    import pyoracle
    meshes, x0s, prm, sc = uneven_blobs()
    s, bos, g = make_pair(oracle, meshes, x0s, prm, sc, precision="mixed")
    for _ in range(3):
        assert pyoracle.world_step(oracle, bos, s) == 0
    g.step(3)
    xs = [x.ravel() * sc for x in x0s]
    for i, (p, bo) in enumerate(zip(g.split(g.positions()), bos)):
        po = bo.get("pos")
        assert np.linalg.norm(p.ravel() - po) <= 1e-4 * np.linalg.norm(po - xs[i]), f"blob {i}"
    assert np.array_equal(g.rng_state(), s.state())


def test_batch_blob_without_force_stays_put():
    """Noise off, every node of blob 1 pinned: its force vector is exactly zero, so its solve
    stops at once with zero velocity (the reference divides 0/0 there), while its neighbours
    relax."""
    from ffea_amd.blob import Batch
    meshes, x0s, prm, sc = uneven_blobs(noise=0)
    meshes[1].pinned = np.arange(meshes[1].num_linear, dtype=np.int32)
    g = Batch(meshes, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), precision="fp64")
    g.set_positions(np.concatenate(x0s) * sc)
    g.step(5)
    its, res, stat = g.last_cg_per_blob()
    assert its[1] == 0 and (stat == 0).all() and its[0] > 0 and its[2] > 0
    p = g.split(g.positions())
    nl = meshes[1].num_linear  # midpoints are re-averaged in scaled units: compare vertices
    assert np.array_equal(p[1][:nl], (x0s[1] * sc)[:nl])
    assert np.linalg.norm(p[0] - x0s[0] * sc) > 0
    assert np.array_equal(g.rng_state(), np.arange(1, 7, dtype=np.uint64))  # noise off: no draws


def test_batch_of_one_equals_blob():
    """A one-blob batch is the same computation as a blob (same plan, same kernels)."""
    from ffea_amd.blob import Batch, Blob
    meshes, x0s, prm, sc = uneven_blobs()
    seed = np.array([12345] * 6, np.uint64)
    a = Blob(meshes[2], prm, sc, rng_state=seed, pos=x0s[2])
    b = Batch([meshes[2]], prm, sc, rng_state=seed)
    b.set_positions(x0s[2] * sc)
    a.step(4)
    b.step(4)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.rng_state(), b.rng_state())
