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
