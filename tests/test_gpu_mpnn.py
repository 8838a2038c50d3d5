"""GPU tests of the learned preconditioner (ffh_mpnn.h) against its CPU restatement
(oracle/mpnn_oracle.py, itself pinned by finite differences), and of the flexible CG that uses
it with the Jacobi fallback.

Bars: the untrained network (zero output layer) is Jacobi bit for bit; forward output within
1e-2 relative of the restatement at the GPU's rounding points (bf16 activations and dense-layer
operands, fp32 MFMA accumulation -- accumulation order is the residual difference); loss within
1e-3 and each parameter block's gradient within 3e-2 relative; learned solves meet the same
||r||^2/||f||^2 < eps^2 as plain ones, and their solutions agree within 1e-4 relative.
"""
import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu


def setup_case(oracle, seed=3):
    import mpnn_oracle as MO
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = box_case((4, 3, 3), noise=0, sigma=0.05, seed=seed, jitter=0.1)
    m.pinned = np.array([0, 7], np.int32)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc, pos=x0)
    assert bo.forces_only(s)[0] == 0
    bo.build_matrix()
    g, lin = MO.graph_from_oracle(bo, m.num_nodes, m.topo, m.pinned)
    b = Blob(m, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), pos=x0)
    assert b.update_internal_forces() == 0
    return m, g, lin, b


def full(m, lin, v):
    out = np.zeros((m.num_nodes, 3))
    out[lin] = v
    return out


def test_untrained_network_is_jacobi(oracle):
    m, g, lin, b = setup_case(oracle)
    b.mpnn_init(seed=1)
    r = np.random.default_rng(0).normal(size=(g.n, 3)) * g.live[:, None]
    z = b.mpnn_apply(full(m, lin, r))
    u = r * g.dinv.reshape(-1, 3)
    assert np.array_equal(z[lin], u)


def test_forward_matches_oracle(oracle):
    import mpnn_oracle as MO
    m, g, lin, b = setup_case(oracle)
    rng = np.random.default_rng(5)
    w = (rng.normal(size=MO.NUM_PARAMS) * 0.3).astype(np.float32)
    b.mpnn_set_weights(w)
    assert np.array_equal(b.mpnn_weights(), w)
    r = rng.normal(size=(g.n, 3)) * g.live[:, None]
    u = r * g.dinv.reshape(-1, 3)
    print("forward: nodes", g.n, flush=True)
    zg = b.mpnn_apply(full(m, lin, r))[lin]
    zo = MO.forward(w, g, u, exact=False)["z"]
    corr_o, corr_g = zo - u, zg - u
    assert np.linalg.norm(corr_g - corr_o) <= 1e-2 * np.linalg.norm(corr_o)


def test_gradient_matches_oracle(oracle):
    import mpnn_oracle as MO
    m, g, lin, b = setup_case(oracle)
    rng = np.random.default_rng(7)
    w = (rng.normal(size=MO.NUM_PARAMS) * 0.3).astype(np.float32)
    b.mpnn_set_weights(w)
    r = rng.normal(size=(g.n, 3)) * g.live[:, None]
    e = rng.normal(size=(g.n, 3)) * g.live[:, None] * np.abs(g.dinv.reshape(-1, 3))
    Lg, Gg = b.mpnn_loss_grad(full(m, lin, r), full(m, lin, e))
    Lo, Go = MO.loss_grad(w, g, r.astype(np.float32), e.astype(np.float32), exact=False)
    assert Lg == pytest.approx(Lo, rel=1e-3)
    po, pg = MO.unpack(Go), MO.unpack(Gg)
    for k in po:
        assert np.linalg.norm(pg[k] - po[k]) <= 3e-2 * np.linalg.norm(po[k]) + 1e-12, k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_training_reduces_loss_and_learned_solve_converges(oracle, precision):
    """Harvest (residual, correction) pairs from plain solves, train, and solve with the
    learned flexible CG: same stopping test, same solution within 1e-4."""
    from ffea_amd.blob import Blob
    m, prm, sc, x0 = box_case((6, 5, 4), noise=1, sigma=0.05, seed=11, jitter=0.1)
    b = Blob(m, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64), pos=x0, precision=precision)
    b.mpnn_init(seed=2)
    b.mpnn_harvest(200)
    b.step(12)
    snaps, pairs = b.mpnn_harvest_count()
    assert snaps > 0 and pairs == snaps
    loss = b.mpnn_train(epochs=8, lr=2e-3, seed=3)
    print("loss per epoch", loss, flush=True)
    assert np.all(np.isfinite(loss)) and loss[-1] < 0.8 * loss[0]
    # the same right-hand side, plain then learned
    assert b.update_internal_forces() == 0
    f = np.random.default_rng(9).normal(size=(m.num_nodes, 3))
    f[m.num_linear:] = 0.0
    b.set_preconditioner("jacobi")
    v_plain = b.solve(f)
    it_plain = b.last_cg()[0]
    b.set_preconditioner("learned")
    v_learn = b.solve(f)
    it_learn, res = b.last_cg()
    print("iterations plain", it_plain, "learned", it_learn, flush=True)
    assert res < prm.epsilon2
    assert np.linalg.norm(v_learn - v_plain) <= 1e-4 * np.linalg.norm(v_plain)


def test_bad_network_falls_back_to_jacobi(oracle):
    """A random (untrained, nonzero output) network for 2 iterations, then the Jacobi restart:
    the solve still meets the stopping test and agrees with the plain solution."""
    import mpnn_oracle as MO
    m, g, lin, b = setup_case(oracle, seed=4)
    w = (np.random.default_rng(3).normal(size=MO.NUM_PARAMS)).astype(np.float32)
    b.mpnn_set_weights(w)
    f = np.random.default_rng(1).normal(size=(m.num_nodes, 3))
    f[m.num_linear:] = 0.0
    f[m.pinned] = 0.0
    v_plain = b.solve(f)
    it_plain = b.last_cg()[0]
    b.set_preconditioner("learned", fallback_iterations=2)
    v = b.solve(f)
    it, res = b.last_cg()
    assert res < b.params.epsilon2 and it > 2
    assert np.linalg.norm(v - v_plain) <= 1e-4 * np.linalg.norm(v_plain)


def test_learned_batch_solve_per_blob():
    """Learned flexible CG on a batch: every blob stops on its own test."""
    from ffea_amd.blob import Batch
    from test_gpu_batch import uneven_blobs
    meshes, x0s, prm, sc = uneven_blobs()
    g = Batch(meshes, prm, sc, rng_state=np.arange(1, 7, dtype=np.uint64))
    g.set_positions(np.concatenate(x0s) * sc)
    g.mpnn_init(seed=5)
    g.mpnn_harvest(100)
    g.step(6)
    loss = g.mpnn_train(epochs=4, lr=2e-3)
    assert np.all(np.isfinite(loss))
    g.set_preconditioner("learned")
    g.step(3)
    its, res, stat = g.last_cg_per_blob()
    assert (stat == 0).all() and (res < prm.epsilon2).all()
