"""The reference's physics pins for the massless path, restated (its checkers are Python 2) and
run on the GPU path with the reference's own fixture meshes and parameters.

- Equipartition (tests/physics/sphere_63_120_nomass/averages.py:51-55): the mean strain energy of
  measurements 40..end of a 1e5-step run (check = 1000) is within 3 % of (3*63 - 6)/2 kT.
- Stokes diffusion (tests/physics/sphere_diffusion/diffusion_nomass.py:49-79): the centroid's mean
  squared displacement between frames `check` steps apart is within 5 % of
  6 kT / (6 pi eta sum(r)) dt check.  The reference analyses frames 5000..10000 of a 1e6-step run;
  here 2000 frames (2e5 steps) after a 500-frame equilibration, same seed and parameters.

Both runs are deterministic (fixed seed, fixed reduction order), so the results do not vary
from run to run.
"""
import os

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
