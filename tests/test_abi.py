"""The C-ABI library loads on a CPU-only host and exports every function include/ffea_hip.h
declares (no compute calls)."""
import os
import re

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ffea_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ffh_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from ffea_amd import _lib
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert L.ffh_version() >= 1


def test_create_rejects_bad_arguments_without_gpu():
    """Argument validation happens before any device call."""
    import numpy as np
    from ffea_amd import _lib
    L = _lib.load()
    z = np.zeros(1)
    h = L.ffh_blob_create(1, 1, z, np.zeros(10, np.int32), z, z, 0, np.zeros(1, np.int32), 0.0, 1.0, 1e-12,
                          10, 0, 0, 1.0, np.zeros(6, np.uint64), 0)
    assert not h
    assert "invalid solver parameters" in _lib.last_error()


def _mesh_args(m, sc=1.0):
    import numpy as np
    return dict(nn=m.num_nodes, ne=m.num_elements,
                pos=np.ascontiguousarray(m.pos * sc, np.float64).ravel(),
                topo=np.ascontiguousarray(m.topo, np.int32).ravel(),
                mat=np.ascontiguousarray(m.material, np.float64).ravel(),
                rad=np.ascontiguousarray(m.stokes_radius * sc, np.float64))


def _create(L, a, pins=None, precision=0, rng=None, eps2=1e-12, maxit=100, dt=0.02):
    import numpy as np
    pins = np.zeros(0, np.int32) if pins is None else np.asarray(pins, np.int32)
    rng = np.arange(1, 7, dtype=np.uint64) if rng is None else rng
    return L.ffh_blob_create(a["nn"], a["ne"], a["pos"], a["topo"], a["mat"], a["rad"], int(pins.size),
                             pins if pins.size else np.zeros(1, np.int32), dt, 1.0, eps2, maxit, 1, 1, 1.0,
                             rng, precision)


def test_mesh_validation_messages_without_gpu():
    """The plan builder's checks (Blob::init's loaders and NoMassCGSolver::init refuse the same
    inputs, src/Blob.cpp:2517-2640, :2880-2947, :3691-3755) run on the host before any device
    call, each with its own message."""
    import numpy as np
    from ffea_amd import _lib, mesh as M
    L = _lib.load()
    m = M.box_mesh(2, 1, 1)
    a = _mesh_args(m)

    def err(**kw):
        h = _create(L, **kw)
        assert not h
        return _lib.last_error()

    bad = dict(a, topo=a["topo"].copy())
    bad["topo"][3] = m.num_nodes + 5
    assert "out of bounds node index" in err(a=bad)
    bad = dict(a, topo=a["topo"].copy())
    bad["topo"][4] = bad["topo"][0]  # a vertex used as an edge midpoint
    assert "both a vertex and an edge midpoint" in err(a=bad)
    bad = dict(a, rad=a["rad"].copy())
    bad["rad"][m.num_linear] = 0.1  # drag on a midpoint
    assert "midpoint node" in err(a=bad)
    bad = dict(a, pos=a["pos"].copy())
    t = m.topo[0, :4]
    for k in (1, 2, 3):  # collapse element 0 onto its first vertex
        bad["pos"][3 * t[k]:3 * t[k] + 3] = bad["pos"][3 * t[0]:3 * t[0] + 3]
    assert "zero rest volume" in err(a=bad)
    assert "pinned node index out of bounds" in err(a=a, pins=[m.num_nodes])
    assert "RngStream state out of range" in err(a=a, rng=np.array([4294967087, 1, 1, 1, 1, 1], np.uint64))
    assert "unknown precision" in err(a=a, precision=7)
    assert "invalid solver parameters" in err(a=a, maxit=0)
    # a valid mesh gets past every check and only then needs a device
    assert "hipStreamCreate" in err(a=a) or "HIP" in err(a=a).upper()


def test_batch_validation_messages_without_gpu():
    """ffh_batch_create: local indices per blob, and no element may join two blobs."""
    import numpy as np
    from ffea_amd import _lib, mesh as M
    L = _lib.load()
    m1, m2 = M.box_mesh(1, 1, 1), M.box_mesh(2, 1, 1)
    nn = np.array([m1.num_nodes, m2.num_nodes], np.int32)
    ne = np.array([m1.num_elements, m2.num_elements], np.int32)
    pos = np.concatenate([m1.pos, m2.pos]).ravel().copy()
    topo = np.concatenate([m1.topo, m2.topo]).astype(np.int32).ravel().copy()
    mat = np.concatenate([m1.material, m2.material]).ravel().copy()
    rad = np.concatenate([m1.stokes_radius, m2.stokes_radius]).copy()
    npin = np.zeros(2, np.int32)

    def create(topo):
        return L.ffh_batch_create(2, nn, ne, pos, topo, mat, rad, npin, np.zeros(1, np.int32), 0.02, 1.0, 1e-12,
                                  100, 1, 1, 1.0, np.arange(1, 7, dtype=np.uint64), 0)
    bad = topo.copy()
    bad[10 * m1.num_elements] = m2.num_nodes  # blob 1's first element: index beyond blob 1
    assert not create(bad) and "topology node index out of bounds" in _lib.last_error()
    assert not create(topo) and "hipStreamCreate" in _lib.last_error()  # valid: stops at the device
    bad = np.array([0, 3], np.int32)
    h = L.ffh_batch_create(2, nn, ne, pos, topo, mat, rad, bad, np.zeros(3, np.int32) + 99, 0.02, 1.0,
                           1e-12, 100, 1, 1, 1.0, np.arange(1, 7, dtype=np.uint64), 0)
    assert not h and "pinned node index out of bounds" in _lib.last_error()
