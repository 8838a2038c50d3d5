"""RngStream restatement (oracle) pinned by the reference's known-answer test, its golden
draws and the algebra of MRG32k3a (jump == repeated single steps)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_rngstream_known_answer(oracle):
    """tests/consistency/rngstream/testRngStream.cpp restated call by call: the sum must print
    as 39.697547445251 at 14 significant digits."""
    o = oracle
    o.set_package_seed(np.array([12345] * 6, np.uint64))
    g1, g2, g3 = o.stream(), o.stream(), o.stream()
    s = g2.randu01() + g3.randu01()
    g1.advance(5, 3)
    s += g1.randu01()
    g1.reset_start_stream()
    for _ in range(35):
        g1.advance(0, 1)
    s += g1.randu01()
    g1.reset_start_stream()
    sumi = 0
    for _ in range(35):
        sumi += g1.randint(1, 10)
    s += sumi / 100.0
    s3 = 0.0
    for _ in range(100):
        s3 += g3.randu01()
    s += s3 / 10.0
    g3.reset_start_stream()
    for _ in range(5):
        s += g3.randu01()
    for _ in range(4):
        g3.reset_next_substream()
    for _ in range(5):
        s += g3.randu01()
    g3.reset_start_substream()
    for _ in range(5):
        s += g3.randu01()
    g2.reset_next_substream()
    s3 = 0.0
    for _ in range(100000):
        s3 += g2.randu01()
    s += s3 / 10000.0
    g3.set_antithetic(True)
    s3 = 0.0
    for _ in range(100000):
        s3 += g3.randu01()
    s += s3 / 10000.0
    assert o.set_package_seed(np.array([1] * 6, np.uint64))
    gar = [o.stream() for _ in range(4)]
    for g in gar:
        s += g.randu01()
    gar[2].advance(-127, 0)
    s += gar[2].randu01()
    gar[2].reset_next_substream()
    gar[2].increased_precis(True)
    s3 = 0.0
    for _ in range(100000):
        s3 += gar[2].randu01()
    s += s3 / 10000.0
    gar[2].set_antithetic(True)
    s3 = 0.0
    for _ in range(100000):
        s3 += gar[2].randu01()
    s += s3 / 10000.0
    gar[2].set_antithetic(False)
    gar[2].increased_precis(False)
    for g in gar:
        s += g.randu01()
    assert "%.14g" % s == "39.697547445251"


def test_package_seed_matches_reference_fixture(oracle):
    from ffea_amd.blob import package_seed
    gold = np.load(os.path.join(GOLDEN, "rng_seed44.npz"))
    assert np.array_equal(oracle.package_seed(44), gold["seed"])
    assert np.array_equal(package_seed(44), gold["seed"])


def _fcp_streams(path):
    with open(path) as fh:
        head = fh.readline()
        n = int(head.split(":")[1])
        return np.array([[int(t) for t in fh.readline().split()] for _ in range(n)], np.uint64)


@pytest.mark.parametrize("who", ["product", "oracle", "reference"])
def test_package_seed_reproduces_reference_checkpoint(oracle, who):
    """The reference's own checkpoint tests/physics/cyl_160_fine/cyl_160_fine.fcp was written by
    an FFEA run of cyl_160_fine-E.ffea (rng_seed = 753, calc_noise = 0, calc_stokes = 0: no draw
    is ever made), with 8 threads, so its 8 lines are the 8 initial RngStreams of World::init
    (src/World.cpp:299-313): package seed from srand(753)/rand(), then each stream 2^127 steps
    after the previous one. All 48 integers must match."""
    from ffea_amd.blob import package_seed
    want = _fcp_streams(os.path.join(GOLDEN, "formats", "cyl_160_fine.fcp"))
    assert want.shape == (8, 6)
    if who == "product":
        seed = package_seed(753)
    elif who == "oracle":
        seed = oracle.package_seed(753)
    else:
        import pyoracle
        if not pyoracle.have_reference():
            pytest.skip("oracle/_ref (the reference's compiled objects) is not built here")
        seed = pyoracle.Reference().package_seed(753)
    assert oracle.set_package_seed(seed)
    got = np.array([oracle.stream().state() for _ in range(8)], np.uint64)
    assert np.array_equal(got, want), (got[0], want[0])


def test_package_seed_does_not_wrap():
    """Seed 44 (every config and the reference's physics fixtures) has rand()+rand() above
    INT_MAX in words 0 and 3: the seed must hold the 64-bit sum, not the wrapped int."""
    import ctypes as C
    from ffea_amd.blob import package_seed
    libc = C.CDLL(None)
    libc.srand(44)
    sums = [int(libc.rand()) + int(libc.rand()) for _ in range(6)]
    assert sums[0] > 2**31 - 1 and sums[3] > 2**31 - 1
    two_max = (4294967087, 4294944443)
    assert [int(v) for v in package_seed(44)] == [s % two_max[i // 3] for i, s in enumerate(sums)]


def test_draws_and_jumps_match_reference_fixture(oracle):
    gold = np.load(os.path.join(GOLDEN, "rng_seed44.npz"))
    oracle.set_package_seed(gold["seed"])
    g = oracle.stream()
    d = np.array([g.randu01() for _ in range(64)])
    assert np.array_equal(d, gold["draws"])
    states = [g.state()]
    for e, c in gold["advances"]:
        g.advance(int(e), int(c))
        states.append(g.state())
    assert np.array_equal(np.array(states), gold["states"])
    assert np.array_equal(np.array([g.randu01() for _ in range(8)]), gold["draws_after"])


@pytest.mark.parametrize("k", [1, 7, 255, 256, 1000, 7 * 3072 + 3 * 4913])
def test_jump_equals_sequential_draws(oracle, k):
    """AdvanceState(0, k) must land where k U01 calls land (brute force)."""
    oracle.set_package_seed(np.array([3015861669, 1834431885, 558792412, 113187832, 1288513475,
                                      2073327934], np.uint64))
    a = oracle.stream()
    oracle.set_package_seed(np.array([3015861669, 1834431885, 558792412, 113187832, 1288513475,
                                      2073327934], np.uint64))
    b = oracle.stream()
    for _ in range(k):
        a.u01()
    b.advance(0, k)
    assert np.array_equal(a.state(), b.state())


def test_jump_two_pow(oracle):
    """AdvanceState(e, 0) == AdvanceState(0, 2^e) for small e."""
    for e in (1, 5, 12):
        oracle.set_package_seed(np.array([1, 2, 3, 4, 5, 6], np.uint64))
        a, b = oracle.stream(), oracle.stream()
        b.set_seed(a.state())
        a.advance(e, 0)
        b.advance(0, 2 ** e)
        assert np.array_equal(a.state(), b.state())


def test_u01_formula_first_draw(oracle):
    """First output from a hand-computed MRG32k3a step (L'Ecuyer 1999 recurrences)."""
    seed = [12345] * 6
    oracle.set_package_seed(np.array(seed, np.uint64))
    g = oracle.stream()
    m1, m2 = 4294967087, 4294944443
    p1 = (1403580 * seed[1] - 810728 * seed[0]) % m1
    p2 = (527612 * seed[5] - 1370589 * seed[3]) % m2
    u = ((p1 - p2) if p1 > p2 else (p1 - p2 + m1)) * (1.0 / (m1 + 1.0))
    assert g.u01() == u
    assert list(g.state()) == [12345, 12345, p1, 12345, 12345, p2]
