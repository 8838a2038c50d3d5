"""The reference's restart test (tests/physics/sphere_63_120_nomass_restart: CMakeLists.txt and
checkRuns.py), restated on the GPU path with its own three scripts: a 10-step run (I), a 5-step
run (II) and its restart to 10 steps from the checkpoint.  The checkpoint files of I and of the
restart must be identical, and the measurements of II (with the restarted steps appended) must
agree with I within 5e-4 relative -- the restart reads positions back from the trajectory's
6-digit text.  Also: the trajectory frames written before the restart are identical, the
restart frame's vertices are the ones read back, later frames agree within 1e-4.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SRC = os.path.join(GOLDEN, "physics")


def measurements(path):
    lines = open(path).read().split("\n")
    k = lines.index("Measurements:")
    return np.array([[float(v) for v in l.split()] for l in lines[k + 2:] if l.strip() and "RESTART" not in l])


def test_restart_reproduces_continuous_run(tmp_path):
    from ffea_amd.world import World, read_trajectory
    shutil.copytree(os.path.join(SRC, "sphere_63_120_structure"), tmp_path / "sphere_63_120_structure")
    run = tmp_path / "restart"
    shutil.copytree(os.path.join(SRC, "sphere_63_120_nomass_restart"), run)
    for name in ("sphere_63_120_nomass_10steps.ffea", "sphere_63_120_nomass_1-5steps.ffea",
                 "sphere_63_120_nomass_6-10steps.ffea"):
        print("run", name, flush=True)
        w = World(str(run / name))
        w.run()
        del w
    a = open(run / "sphere_63_120_nomass_6-10steps.fcp").read().split("\n")
    b = open(run / "sphere_63_120_nomass_10steps.fcp").read().split("\n")
    assert a[:2] == b[:2], "checkpoint files should be exactly the same"
    m1 = measurements(run / "sphere_63_120_nomass_I_measurement.out")
    m2 = measurements(run / "sphere_63_120_nomass_II_measurement.out")
    assert m1.shape == m2.shape == (11, 7)  # SurfSurfEnergy too: calc_steric is on by default
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(m1 != 0, np.abs((m2 - m1) / m1), 0.0)
    assert d.max() <= 5e-4, d.max()
    f1 = read_trajectory(str(run / "sphere_63_120_nomass_I_trajectory.out"))
    f2 = read_trajectory(str(run / "sphere_63_120_nomass_II_trajectory.out"))
    assert [f[0] for f in f1] == [f[0] for f in f2] == list(range(11))
    from ffea_amd import formats as F
    topo, _ = F.read_topology(os.path.join(SRC, "sphere_63_120_structure", "sphere_63_120.top"))
    lin = np.unique(topo[:, :4])  # the vertices; the other nodes are edge midpoints
    assert lin.size == 63
    for (s1, b1, _), (s2, b2, _) in zip(f1, f2):
        if s1 < 5:  # written before the restart
            assert np.array_equal(b1[0], b2[0])
        elif s1 == 5:  # rewritten from the frame read back: vertices as read, midpoints re-averaged
            assert np.array_equal(b1[0][lin], b2[0][lin])
        assert np.allclose(b2[0], b1[0], rtol=1e-4, atol=1e-4 * np.abs(b1[0]).max())


def test_mass_restart_reproduces_continuous_run(tmp_path):
    """The same three-run restart check with the mass-ful solver (solver = CG): velocities are
    read back from the trajectory as well (Blob::read_nodes_from_file, src/Blob.cpp:1289-1301);
    the kinetic energy joins the measurements."""
    from ffea_amd.world import World
    shutil.copytree(os.path.join(SRC, "sphere_63_120_structure"), tmp_path / "sphere_63_120_structure")
    run = tmp_path / "restart"
    shutil.copytree(os.path.join(SRC, "sphere_63_120_nomass_restart"), run)
    names = ("sphere_63_120_nomass_10steps.ffea", "sphere_63_120_nomass_1-5steps.ffea",
             "sphere_63_120_nomass_6-10steps.ffea")
    for name in names:
        p = run / name
        p.write_text(p.read_text().replace("<solver = CG_nomass>", "<solver = CG>"))
    for name in names:
        print("run", name, flush=True)
        w = World(str(run / name))
        assert w.mass_in_system
        w.run()
        del w
    a = open(run / "sphere_63_120_nomass_6-10steps.fcp").read().split("\n")
    b = open(run / "sphere_63_120_nomass_10steps.fcp").read().split("\n")
    assert a[:2] == b[:2]
    m1 = measurements(run / "sphere_63_120_nomass_I_measurement.out")
    m2 = measurements(run / "sphere_63_120_nomass_II_measurement.out")
    assert m1.shape == m2.shape == (11, 8)  # kinetic energy and SurfSurfEnergy too
    assert m1[1:, 1].min() > 0  # kinetic energy
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(m1 != 0, np.abs((m2 - m1) / m1), 0.0)
    assert d.max() <= 5e-4, d.max()


def test_world_wraps_blob_into_periodic_box(tmp_path):
    """World::run's periodic box through the World driver: the sphere placed at centroid
    (-2, 0.3, -1.5) (times its 5 nm scale) without move_into_box; the box is 2 ceil(extent /
    3 nm) cells of 3 nm per axis (World::init, src/World.cpp:455-471).  At step 0 the axes whose
    centroid lies below zero move up by one box length (src/World.cpp:1850-1901), so the first
    frame's centroid is inside the box; the other axis is untouched."""
    from ffea_amd import units as U
    from ffea_amd.world import World, read_trajectory
    shutil.copytree(os.path.join(SRC, "sphere_63_120_structure"), tmp_path / "sphere_63_120_structure")
    run = tmp_path / "box"
    shutil.copytree(os.path.join(SRC, "sphere_63_120_nomass"), run)
    p = run / "sphere_63_120_nomass.ffea"
    t = p.read_text().replace("<num_steps = 1e5>", "<num_steps = 3>").replace("<check = 1e3>", "<check = 1>")
    t = t.replace("<centroid = (0.0,0.0,0.0)>", "<centroid = (-2.0,0.3,-1.5)>")
    t = t.replace("</param>", "\t<move_into_box = 0>\n</param>")
    p.write_text(t)
    w = World(str(p))
    assert w.walls == 0b111111
    box = w.box * U.MESO_LENGTH
    w.run()
    frames = read_trajectory(str(run / "sphere_63_120_nomass_trajectory.out"))
    c0 = frames[0][1][0][:, :3].mean(axis=0)
    placed = np.array([-2.0, 0.3, -1.5]) * 5e-9
    assert np.allclose(c0, placed + np.array([box[0], 0.0, box[2]]), rtol=0, atol=1e-3 * box.max())
    assert (c0 > 0).all() and (c0 < box).all()
