"""Live bitwise comparison of the oracle with the reference's compiled objects
(oracle/_ref) on seeded random meshes; skipped where oracle/_ref is not built."""
import numpy as np
import pytest

import pyoracle
from conftest import box_case


@pytest.mark.parametrize("case", [((2, 2, 2), 0.0, 11, 1), ((4, 3, 2), 0.15, 12, 1), ((3, 3, 3), 0.1, 13, 0)])
def test_trajectory_bitwise(oracle, reference, case):
    n, jit, seed, noise = case
    m, prm, sc, x0 = box_case(n, noise=noise, jitter=jit, seed=seed)
    m.pinned = np.array([seed % 7, 2 * seed % 11], np.int32)
    sd = oracle.package_seed(seed)
    oracle.set_package_seed(sd)
    reference.set_package_seed(sd)
    so, sr = oracle.stream(), reference.stream()
    bo, br = oracle.blob(m, prm, sc, pos=x0), reference.blob(m, prm, sc, pos=x0)
    for _ in range(6):
        assert pyoracle.world_step(oracle, [bo], so) == pyoracle.world_step(reference, [br], sr) == 0
        for k in ("pos", "vel", "node_force", "visc", "F_ij", "stress_mag"):
            assert np.array_equal(bo.get(k), br.get(k)), k
    assert np.array_equal(so.state(), sr.state())
    ko, co, vo = bo.csr()
    kr, cr, vr = br.csr()
    assert np.array_equal(ko, kr) and np.array_equal(co, cr) and np.array_equal(vo, vr)
    x = np.random.default_rng(seed).normal(size=3 * m.num_nodes)
    assert np.array_equal(bo.apply_matrix(x), br.apply_matrix(x))


def test_two_blobs_share_one_stream(oracle, reference):
    """World::run order over several blobs: all element phases, then all position phases."""
    cases = [box_case((2, 2, 2), noise=1, seed=21), box_case((3, 2, 2), noise=1, jitter=0.1, seed=22)]
    sd = oracle.package_seed(5)
    oracle.set_package_seed(sd)
    reference.set_package_seed(sd)
    so, sr = oracle.stream(), reference.stream()
    bo = [oracle.blob(m, p, s, pos=x) for m, p, s, x in cases]
    br = [reference.blob(m, p, s, pos=x) for m, p, s, x in cases]
    for _ in range(3):
        assert pyoracle.world_step(oracle, bo, so) == 0
        assert pyoracle.world_step(reference, br, sr) == 0
    for a, b in zip(bo, br):
        assert np.array_equal(a.get("pos"), b.get("pos"))
    assert np.array_equal(so.state(), sr.state())
