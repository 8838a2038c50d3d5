import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SPHERE_FFEA = os.path.join(GOLDEN, "physics", "sphere_63_120_nomass", "sphere_63_120_nomass.ffea")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libffea_hip.so)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build(with_ref=False)
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    import pyoracle
    if not pyoracle.have_reference():
        pytest.skip("oracle/_ref (the reference's compiled objects) is not built here")
    return pyoracle.Reference()


def sphere_case():
    """The reference's own sphere_63_120 nomass fixture: mesh, params, scale."""
    from ffea_amd import formats as F, units as U
    s = F.read_script(SPHERE_FFEA)
    b = s["blobs"][0]
    m = F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin"))
    return m, F.params_from_script(s["params"]), U.blob_scale(float(b["scale"]))


def box_case(n=(3, 3, 2), noise=1, eps=1e-6, sigma=0.02, seed=1, jitter=0.0):
    from ffea_amd import mesh as M, units as U
    m = M.box_mesh(*n, jitter=jitter, seed=seed)
    prm = U.SimParams.from_si(epsilon=eps, calc_noise=noise)
    x0 = M.perturb(m, sigma, seed=seed)
    return m, prm, U.blob_scale(1e-9), x0
