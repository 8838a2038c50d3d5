"""The learned preconditioner on the c3 workload's 64 held-out meshes (jittered boxes with n in
[24, 64] per axis, 34.5 M tets, mixed precision, bench.py's make_case("c3")).

The network is trained on the GPU on (residual, correction) pairs harvested from plain CG solves
of a smaller batch drawn like c2 (training data never includes the held-out meshes).  From one
state of the held-out batch, one step each with:
  - Jacobi-PCG (the fallback and the judge, src/NoMassCGSolver.cpp:278-292);
  - the untrained network (zero output layer: its output is exactly Jacobi's, inside the
    flexible CG's Polak-Ribiere beta, so within one iteration of Jacobi-PCG per blob);
  - the trained network inside flexible CG.
Every blob meets the same stopping test, the learned velocities agree with Jacobi's within 1e-4
per blob, and the trained network needs fewer iterations than the untrained one / Jacobi, in
total over the 64 blobs and on most blobs: the gain that training buys.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_c3_learned_against_jacobi_on_held_out_meshes():
    import bench
    from ffea_amd.blob import Batch, package_seed
    ms, prm, sc, xs, _ = bench.make_case("c3")
    # training batch: 24 c2-style blobs (n in [6, 32]), seeds distinct from the held-out set
    tm, tx = bench.box_batch(24, 6, 32, seed=13)
    tb = Batch(tm, prm, sc, rng_state=package_seed(prm.rng_seed), precision="mixed")
    tb.set_positions(np.concatenate(tx) * sc)
    tb.mpnn_init(seed=1)
    tb.mpnn_harvest(600)
    steps = 0
    while tb.mpnn_harvest_count()[1] < 6000 and steps < 100:
        tb.step(1)
        steps += 1
    snaps, pairs = tb.mpnn_harvest_count()
    loss = tb.mpnn_train(epochs=12, lr=2e-3, seed=1)
    print(f"harvested {pairs} pairs in {steps} steps; loss {loss[0]:.4f} -> {loss[-1]:.4f}", flush=True)
    assert np.all(np.isfinite(loss)) and loss[-1] < loss[0]
    w = tb.mpnn_weights()
    tb.close()

    g = Batch(ms, prm, sc, rng_state=package_seed(prm.rng_seed), precision="mixed")
    x0 = np.concatenate(xs) * sc
    print("held-out batch", g.sizes(), flush=True)
    st0 = None
    runs = {}
    for name in ("jacobi", "untrained", "trained"):
        g.set_positions(x0)
        if st0 is None:
            st0 = g.rng_state()
        g.set_rng_state(st0)
        if name == "jacobi":
            g.set_preconditioner("jacobi")
        else:
            g.mpnn_init(seed=2) if name == "untrained" else g.mpnn_set_weights(w)
            g.set_preconditioner("jacobi")
            g.step(1)  # a plain solve sets the fallback count (the last plain iteration count)
            g.set_positions(x0)
            g.set_rng_state(st0)
            g.set_preconditioner("learned")
        g.step(1)
        its, res, stat = g.last_cg_per_blob()
        assert (stat == 0).all() and (res < prm.epsilon2).all(), name
        runs[name] = (its.copy(), [v.copy() for v in g.split(g.velocities())])
        print(name, "iterations per blob: total", its.sum(), "mean", its.mean(), "max", its.max(), flush=True)
    jac, unt, trn = runs["jacobi"][0], runs["untrained"][0], runs["trained"][0]
    assert np.abs(unt.astype(int) - jac).max() <= 1  # Jacobi, under the flexible CG's beta
    for name in ("untrained", "trained"):
        for k, (a, b) in enumerate(zip(runs[name][1], runs["jacobi"][1])):
            assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b), (name, k)
    assert trn.sum() < jac.sum() and trn.sum() < unt.sum()
    assert (trn < jac).sum() > len(ms) // 2
