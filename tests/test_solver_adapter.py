"""The Solver adapter (ffea_amd/integration/GpuNoMassCGSolver: include/Solver.h:30-55 over the
C-ABI), compiled against the reference's own headers (oracle/_ref/libffea_adapter.so) and driven
on a reference blob's own mesh_node / tetra_element_linear / SimulationParams objects, built by
the reference harness as Blob::init builds them.

CPU: init's marshalling of those objects (positions, node indices, meso materials, radii, the
pinned and binding-site pinned nodes together) and its error paths.
GPU: init / solve / apply_matrix against the reference's own NoMassCGSolver on the same
objects -- the matrix bit for bit (the reference CSR's sequential row sums), solutions within
1e-9 (CG dot products in tree order), also after the positions move; a CG failure is reported
as FFEA_ERROR.
"""
import os

import numpy as np
import pytest

from conftest import box_case


def adapter_case(pinned=(), radius_on_midpoint=False, **kw):
    import pyoracle
    if not os.path.exists(pyoracle.ADAPTER_SO) or not pyoracle.have_reference():
        pytest.skip("oracle/_ref (the reference's objects and the adapter build) is not built here")
    m, prm, sc, x0 = box_case((3, 3, 2), noise=1, jitter=0.1, seed=4, **kw)
    m.pinned = np.array(pinned, np.int32)
    if radius_on_midpoint:
        m.stokes_radius[m.num_linear + 2] = 0.5
    ref = pyoracle.Reference()
    ref.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    bo = ref.blob(m, prm, sc, pos=x0)
    return ref, bo, pyoracle.GpuSolverAdapter(ref, bo, m.num_nodes, m.num_elements), m, prm, sc, x0


def test_marshalling_of_reference_objects():
    ref, bo, ad, m, prm, sc, x0 = adapter_case(pinned=(0, 5))
    a = ad.marshal(pinned=[5, 0], bsite=[7, 5, 12])
    assert np.array_equal(a["pos"], bo.get("pos"))
    assert np.array_equal(a["topo"], m.topo.ravel())
    assert np.array_equal(a["material"], bo.get("material"))  # rho, A, B, G, E as the elements hold them
    assert np.array_equal(a["radius"], m.stokes_radius * sc)
    assert list(a["pinned"]) == [0, 5, 7, 12]
    with pytest.raises(ValueError, match="out of range"):
        ad.marshal(pinned=[m.num_nodes])
    ad.close()
    bo.close()


def test_init_refuses_midpoint_drag_with_the_librarys_message():
    """Midpoint rows must carry no Stokes drag on the GPU path (the rows are dropped); the
    reference accepts such a mesh, the adapter's init returns FFEA_ERROR with the reason."""
    from ffea_amd import _lib
    ref, bo, ad, m, prm, sc, x0 = adapter_case(radius_on_midpoint=True)
    assert ad.init() == -1  # FFEA_ERROR
    assert "nonzero Stokes radius" in _lib.last_error()
    ad.close()
    bo.close()


@pytest.mark.gpu
def test_solve_and_apply_matrix_against_reference_solver():
    import pyoracle
    from test_gpu_parity import seq_csr_apply
    ref, bo, ad, m, prm, sc, x0 = adapter_case(pinned=(0, 3))
    assert ad.init(pinned=m.pinned) == 0  # the blob's pinned_nodes_list, as Blob::init passes it
    s = ref.stream()
    for k in range(2):
        inv, rhs = bo.forces_only(s)  # element loop + force gather: the element matrices at these positions
        assert inv == 0
        st, v_ref = bo.solve(rhs)
        assert st == 0
        rc, v = ad.solve(rhs)
        assert rc == 0
        assert np.linalg.norm(v - v_ref) <= 1e-9 * np.linalg.norm(v_ref), k
        key, col, val = bo.csr()
        x = np.random.default_rng(k).normal(size=3 * m.num_nodes)
        assert np.array_equal(ad.apply_matrix(x), seq_csr_apply(key, col, val, x)), k
        assert pyoracle.world_step(ref, [bo], s) == 0  # the blob moves on; the adapter re-assembles there
    ad.close()
    bo.close()


@pytest.mark.gpu
def test_cg_failure_is_ffea_error():
    ref, bo, ad, m, prm, sc, x0 = adapter_case(eps=1e-100)
    import ctypes as C
    # one CG iteration allowed: the reference's SimulationParams field as the blob holds it
    ref.lib.ref_blob_set_max_iterations_cg.argtypes = [C.c_void_p, C.c_int]
    ref.lib.ref_blob_set_max_iterations_cg(bo.h, 1)
    assert ad.init() == 0
    inv, rhs = bo.forces_only(ref.stream())
    rc, v = ad.solve(rhs)
    assert rc == -1
    ad.close()
    bo.close()
