"""The reference's physics pins, restated (its checkers are Python 2) and run on the GPU path
with the reference's own fixture meshes and parameters.

- Equipartition (tests/physics/sphere_63_120_nomass/averages.py:51-55): the mean strain energy of
  measurements 40..end of a 1e5-step run (check = 1000) is within 3 % of (3*63 - 6)/2 kT.
- Stokes diffusion (tests/physics/sphere_diffusion/diffusion_nomass.py:49-79): the centroid's mean
  squared displacement between frames `check` steps apart is within 5 % of
  6 kT / (6 pi eta sum(r)) dt check.  The reference analyses frames 5000..10000 of a 1e6-step run;
  here 2000 frames (2e5 steps) after a 500-frame equilibration, same seed and parameters.

- Mass-ful equipartition (tests/physics/sphere_63_120_mass/averages.py:51-74), through the
  World driver and its measurement file: with solver = CG, measurements 40..end of a 1e5-step run
  give the kinetic energy within 4 % of 3*63/2 kT and the strain energy within 3 % of
  (3*63 - 6)/2 kT.

- Inertial diffusion (tests/physics/sphere_diffusion/diffusion_mass.py:49-79): with solver = CG
  the centroid's mean squared displacement between frames `check` steps apart against
  3 kT / M (dt check)^2.  The reference analyses frames 5000..10000 of a 2e6-step run; its own
  objects' restatement (the oracle) gives a relative error of 0.086 there (bar 0.10).  Here 1000
  frames after 250 frames of equilibration (2.5e5 steps), where the oracle gives 0.102: the GPU
  run must reproduce the oracle's mean squared displacement within 1e-6 relative, and the
  theory within 0.15.

- Young's modulus and flexural rigidity of a cylinder under surface forces
  (tests/physics/E_cyl_160_fine/calcE.py:42-93, tests/physics/EI_cyl_160_fine/calcEI.py:43-97):
  the reference's cyl_160_fine mesh, one end pinned, a linear surface force on the other end's
  faces (ctforces CTF 3), run through the World driver with the reference's scripts; from the
  end nodes' centroid displacement E = F L / (pi r^2 dL) within 0.5 % of 6e8 Pa and
  EI = F L^3 / (3 dL) within 10 % of E pi r^4 / 4.

The runs are deterministic (fixed seed, fixed reduction order), so the results do not vary
from run to run.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, SPHERE_FFEA

pytestmark = pytest.mark.gpu


def load(script):
    from ffea_amd import formats as F, units as U
    s = F.read_script(script)
    b = s["blobs"][0]
    m = F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin"))
    raw = dict(s["params"], scale=b["scale"])
    return m, F.params_from_script(s["params"]), U.blob_scale(float(b["scale"])), raw


def test_equipartition_sphere_63_120():
    from ffea_amd.blob import Blob, package_seed
    from ffea_amd.units import MESO_ENERGY
    m, prm, sc, raw = load(SPHERE_FFEA)
    b = Blob(m, prm, sc, rng_state=package_seed(prm.rng_seed), precision="fp64")
    check, num_steps = 1000, 100000
    energies = []
    for s in range(0, num_steps + 1, check):
        # the reference measures after the element phase of step s, i.e. at the positions the
        # step starts from (src/World.cpp:2024-2034)
        energies.append(b.strain_energy())
        if s < num_steps:
            b.step(check)
    kT = float(raw["kT"])
    theory = kT * (3 * 63 - 6) / 2
    mean = np.mean(np.array(energies[40:-1]) * MESO_ENERGY)
    err = abs(mean / theory - 1)
    print(f"equipartition: <U> = {mean:.4e} J, theory {theory:.4e} J, rel err {err:.4f}")
    assert err < 0.03


def test_stokes_diffusion_sphere():
    from ffea_amd.blob import Blob, package_seed
    from ffea_amd.units import MESO_LENGTH
    script = os.path.join(GOLDEN, "physics", "sphere_diffusion", "sphere_nomass_norestart.ffea")
    m, prm, sc, raw = load(script)
    b = Blob(m, prm, sc, rng_state=package_seed(prm.rng_seed), precision="fp64")
    check = int(float(raw["check"]))
    b.step(500 * check)
    cents = []
    for _ in range(2001):
        cents.append(b.positions().mean(axis=0) * MESO_LENGTH)  # FFEA_frame.calc_centroid: all nodes
        b.step(check)
    d = np.diff(np.array(cents), axis=0)
    r2mean = np.mean(np.sum(d * d, axis=1))
    drag = 6 * np.pi * float(raw["stokes_visc"]) * m.stokes_radius.sum() * float(raw["scale"])
    r2theory = 6 * (float(raw["kT"]) / drag) * float(raw["dt"]) * check
    err = abs(r2mean - r2theory) / r2theory
    print(f"diffusion: <dr^2> = {r2mean:.4e} m^2, theory {r2theory:.4e} m^2, rel err {err:.4f}")
    assert err <= 0.05


def fm_columns(path):
    from ffea_amd.analysis import read_measurements
    return read_measurements(path)


def test_mass_equipartition_sphere_63_120(tmp_path):
    from ffea_amd.world import World
    shutil.copytree(os.path.join(GOLDEN, "physics", "sphere_63_120_structure"), tmp_path / "sphere_63_120_structure")
    run = tmp_path / "sphere_63_120_mass"
    shutil.copytree(os.path.join(GOLDEN, "physics", "sphere_63_120_mass"), run)
    w = World(str(run / "sphere_63_120_mass.ffea"))
    assert w.solver == "CG" and w.mass_in_system
    w.run()
    from ffea_amd.analysis import equipartition
    H = fm_columns(run / "sphere_63_120_mass_measurement.out")
    errs = equipartition(H, nodes=63, kT=4.11e-21, start=39)  # the checker's H rows start with the name
    tol = {"KineticEnergy": 0.04, "StrainEnergy": 0.03}
    for k in tol:
        print(f"{k}: rel err {errs[k]:.4f} (tolerance {tol[k]})", flush=True)
        assert errs[k] < tol[k], k


def test_inertial_diffusion_sphere(oracle):
    import pyoracle
    from ffea_amd.blob import Blob, package_seed
    script = os.path.join(GOLDEN, "physics", "sphere_diffusion", "sphere_mass_norestart.ffea")
    m, prm, sc, raw = load(script)
    check, equil, frames = int(float(raw["check"])), 250, 1000
    b = Blob(m, prm, sc, rng_state=package_seed(prm.rng_seed), precision="fp64")
    b.set_solver("CG")
    b.step(equil * check)
    cents = []
    for _ in range(frames + 1):
        cents.append(b.positions().mean(axis=0))  # FFEA_frame.calc_centroid: all nodes
        b.step(check)
    d = np.diff(np.array(cents), axis=0)
    r2 = np.mean(np.sum(d * d, axis=1))
    # the oracle over the same steps
    oracle.set_package_seed(package_seed(prm.rng_seed))
    s = oracle.stream()
    bo = oracle.blob(m, prm, sc)
    bo.set_solver("CG")
    for _ in range(equil * check):
        pyoracle.world_step(oracle, [bo], s)
    co = []
    for _ in range(frames + 1):
        co.append(bo.get("pos").reshape(-1, 3).mean(axis=0))
        for _ in range(check):
            pyoracle.world_step(oracle, [bo], s)
    do = np.diff(np.array(co), axis=0)
    r2o = np.mean(np.sum(do * do, axis=1))
    mass = (bo.get("material")[0::5] * bo.get("vol_0")).sum()
    theory = 3 * (prm.kT / mass) * (prm.dt * check) ** 2
    err = abs(r2 - theory) / theory
    print(f"inertial diffusion: <dr^2> = {r2:.6e}, oracle {r2o:.6e}, theory {theory:.6e}, rel err {err:.4f}")
    assert abs(r2 - r2o) <= 1e-6 * r2o
    assert err <= 0.15


# the end nodes both checkers average over (calcE.py:39, calcEI.py:40)
CYL_ENDNODES = [60, 61, 63, 66, 67, 69, 72, 73, 75, 78, 79, 81, 84, 85, 87, 90, 91, 93, 96, 97, 99, 102, 103, 105,
                108, 109, 111, 114, 116, 156, 160, 164, 168, 172, 176, 180, 184, 188] + list(range(4310, 4402))


def cylinder_run(tmp_path, kind):
    """Rebuilds the cyl_160_fine files from the fixture arrays and runs the reference's script
    through the World driver; returns the end nodes' centroid displacement between the first
    and the last frame (m)."""
    from ffea_amd import formats as F
    from ffea_amd.mesh import Mesh
    from ffea_amd.world import World, read_trajectory
    z = np.load(os.path.join(GOLDEN, "physics", "cyl_160_fine.npz"))
    d = tmp_path / "cyl_160_fine"
    d.mkdir()
    m = Mesh(pos=z["pos"], topo=z["topo"], material=z["material"], stokes_radius=z["stokes"],
             num_linear=int(np.unique(z["topo"][:, :4]).size), pinned=z["pinned"])
    F.write_blob(str(d / "cyl_160_fine"), m)
    F.write_surface(str(d / "cyl_160_fine.surf"), z["surface"])
    F.write_ssint_types(str(d / "cyl_160_fine.vdw"), -np.ones(z["surface"].shape[0], np.int32))  # all off, as there
    run = tmp_path / f"{kind}_cyl_160_fine"
    shutil.copytree(os.path.join(GOLDEN, "physics", f"cyl_160_fine_{kind}"), run)
    w = World(str(run / f"cyl_160_fine-{kind}.ffea"))
    assert w.calc_ctforces
    w.run()
    from ffea_amd.analysis import end_displacement
    path = str(run / "cyl_160_fine_trajectory.ftj")
    return end_displacement(path, CYL_ENDNODES), len(read_trajectory(path))


def test_youngs_modulus_cylinder(tmp_path):
    from ffea_amd.analysis import youngs_modulus
    dL, nf = cylinder_run(tmp_path, "E")
    F, L, r, E_t = 1e-11, 160e-9, 10e-9, 6e8
    E_c = youngs_modulus(dL, F, L, r)
    print(f"Young's modulus: E_c = {E_c:.5e} Pa over {nf} frames, E_c/E_t = {E_c / E_t:.5f}", flush=True)
    assert abs(E_c / E_t - 1) <= 0.005


def test_flexural_rigidity_cylinder(tmp_path):
    from ffea_amd.analysis import flexural_rigidity
    dL, nf = cylinder_run(tmp_path, "EI")
    F, L, r, E_t = 1e-12, 160e-9, 10e-9, 6e8
    EI_c, EI_t = flexural_rigidity(dL, F, L), E_t * np.pi * r ** 4 / 4.0
    print(f"flexural rigidity: EI_c/EI_t = {EI_c / EI_t:.5f} over {nf} frames", flush=True)
    assert abs(EI_c / EI_t - 1) <= 0.1


def test_cubes_lennard_jones_energy(tmp_path):
    """The reference's tests/physics/cubes_lj through the World driver: two unit cubes whose top
    faces face each other 1 nm apart; checkE.py wants the step-0 SurfSurfEnergy within 1 % of
    the double surface integral -1.7870e-19 J (computed with CUBA)."""
    from ffea_amd.world import World
    run = tmp_path / "cubes_lj"
    shutil.copytree(os.path.join(GOLDEN, "physics", "cubes_lj"), run)
    w = World(str(run / "cube_veryFine.ffea"))
    w.run()
    from ffea_amd.analysis import surface_energy
    e = surface_energy(fm_columns(run / "veryFinefine1face1face.fm"))
    print(f"cubes_lj: SurfSurfEnergy {e:.6e} J, rel err {abs(e / -1.7870e-19 - 1):.4f}", flush=True)
    assert abs((e - -1.7870e-19) / -1.7870e-19) < 1e-2
