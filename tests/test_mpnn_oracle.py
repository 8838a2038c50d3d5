"""Pins of the learned preconditioner's CPU restatement (oracle/mpnn_oracle.py): its backward
pass against central finite differences of its loss, the zero output layer as exact Jacobi,
and positive homogeneity (z(a r) = a z(r), a > 0, from the per-blob scale)."""
import numpy as np
import pytest

from conftest import box_case


@pytest.fixture(scope="module")
def graph(oracle):
    import mpnn_oracle as MO
    m, prm, sc, x0 = box_case((2, 2, 2), noise=0, sigma=0.05, seed=3, jitter=0.1)
    m.pinned = np.array([0, 5], np.int32)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    s = oracle.stream()
    b = oracle.blob(m, prm, sc, pos=x0)
    assert b.forces_only(s)[0] == 0
    b.build_matrix()
    g, lin = MO.graph_from_oracle(b, m.num_nodes, m.topo, m.pinned)
    return g


def rand_pair(g, seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=(g.n, 3)) * g.live[:, None]
    e = rng.normal(size=(g.n, 3)) * g.live[:, None]
    return r, e


def test_backward_matches_finite_differences(graph):
    import mpnn_oracle as MO
    rng = np.random.default_rng(0)
    w = rng.normal(size=MO.NUM_PARAMS) * 0.3
    r, e = rand_pair(graph, 1)
    L, G = MO.loss_grad(w, graph, r, e, exact=True)
    assert L > 0
    idx = np.concatenate([rng.choice(MO.NUM_PARAMS, 60, replace=False), np.arange(MO.OFF_WO, MO.NUM_PARAMS),
                          np.arange(MO.OFF_BE, MO.OFF_LAY)])
    gmax = np.abs(G).max()
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (MO.loss(wp, graph, r, e) - MO.loss(wm, graph, r, e)) / (2 * h)
        assert G[i] == pytest.approx(fd, rel=1e-4, abs=1e-6 * gmax), i


def test_zero_output_layer_is_jacobi(graph):
    import mpnn_oracle as MO
    w = np.random.default_rng(2).normal(size=MO.NUM_PARAMS)
    w[MO.OFF_WO:] = 0.0
    r, _ = rand_pair(graph, 3)
    u = r * graph.dinv.reshape(-1, 3)
    for exact in (True, False):
        z = MO.forward(w, graph, u, exact)["z"]
        assert np.array_equal(z[graph.live], u[graph.live])
        assert np.all(z[~graph.live] == 0)


def test_positive_homogeneity(graph):
    import mpnn_oracle as MO
    w = np.random.default_rng(4).normal(size=MO.NUM_PARAMS) * 0.3
    r, _ = rand_pair(graph, 5)
    u = r * graph.dinv.reshape(-1, 3)
    z1 = MO.forward(w, graph, u, True)["z"]
    z2 = MO.forward(w, graph, 8.0 * u, True)["z"]
    assert np.allclose(z2, 8.0 * z1, rtol=1e-12, atol=1e-14 * np.abs(z2).max())
