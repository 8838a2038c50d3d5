"""The reference's physics checkers as Python 3 functions over this package's output files
(its own scripts under tests/physics are Python 2 and read through ffeatools).

Each function states the checker it restates.  They read the measurement (.fm) and trajectory
(.ftj) files ffea_amd.world writes, in the reference's formats, so they apply to the
reference's own output as well.
"""
from __future__ import annotations

import numpy as np

from .world import read_trajectory

KB = 1.380649e-23


def read_measurements(path) -> dict:
    """readAndAverage's parse of a measurement file (tests/physics/sphere_63_120_mass/
    averages.py:27-41): {column: values} after the "Measurements:" header; RESTART lines
    skipped."""
    lines = open(path).read().split("\n")
    k = lines.index("Measurements:")
    names = lines[k + 1].split()
    rows = [[float(v) for v in l.split()] for l in lines[k + 2:] if l.strip() and "RESTART" not in l]
    return {n: np.array([r[i] for r in rows]) for i, n in enumerate(names)}


def equipartition(measurements: dict, nodes: int, kT: float, start: int = 40, end: int = -1) -> dict:
    """tests/physics/sphere_63_120_nomass/averages.py:44-60 and sphere_63_120_mass/averages.py:
    relative deviation of the mean strain energy over measurements [start, end) from
    (3 nodes - 6) kT / 2, and of the mean kinetic energy (when measured) from 3 nodes kT / 2."""
    theory = {"StrainEnergy": kT * (3 * nodes - 6) / 2, "KineticEnergy": kT * (3 * nodes) / 2}
    return {k: abs(np.mean(measurements[k][start:end]) / th - 1) for k, th in theory.items() if k in measurements}


def frame_centroids(traj_path, blob: int = 0) -> np.ndarray:
    """FFEA_frame.calc_centroid of every complete frame: the mean of all nodes (m)."""
    return np.array([f[1][blob][:, :3].mean(axis=0) for f in read_trajectory(traj_path)])


def stokes_diffusion(centroids, kT: float, stokes_visc: float, radius_sum: float, scale: float, dt: float,
                     check: int):
    """tests/physics/sphere_diffusion/diffusion_nomass.py:49-79: the mean squared displacement
    of the centroid between consecutive frames against 6 kT / (6 pi eta sum(r) scale) dt check.
    Returns (measured, theory, relative error)."""
    d = np.diff(np.asarray(centroids), axis=0)
    r2 = float(np.mean(np.sum(d * d, axis=1)))
    drag = 6 * np.pi * stokes_visc * radius_sum * scale
    theory = 6 * (kT / drag) * dt * check
    return r2, theory, abs(r2 - theory) / theory


def inertial_diffusion(centroids, kT: float, mass: float, dt: float, check: int):
    """tests/physics/sphere_diffusion/diffusion_mass.py:49-79: the centroid's mean squared
    displacement between consecutive frames against 3 kT / M (dt check)^2 (free flight).
    Returns (measured, theory, relative error)."""
    d = np.diff(np.asarray(centroids), axis=0)
    r2 = float(np.mean(np.sum(d * d, axis=1)))
    theory = 3 * (kT / mass) * (dt * check) ** 2
    return r2, theory, abs(r2 - theory) / theory


def end_displacement(traj_path, end_nodes, blob: int = 0) -> float:
    """The displacement of the end nodes' centroid between the first and the last frame
    (tests/physics/E_cyl_160_fine/calcE.py:68-89, EI_cyl_160_fine/calcEI.py:63-78), m."""
    frames = read_trajectory(traj_path)
    cm = [f[1][blob][list(end_nodes), :3].mean(axis=0) for f in (frames[0], frames[-1])]
    return float(np.sqrt(((cm[1] - cm[0]) ** 2).sum()))


def youngs_modulus(dL: float, F: float, L: float, r: float) -> float:
    """calcE.py:91: E = F L / (pi r^2 dL) for a cylinder of length L and radius r pulled by F."""
    return F * L / (np.pi * r ** 2 * dL)


def flexural_rigidity(dL: float, F: float, L: float) -> float:
    """calcEI.py:81: EI = F L^3 / (3 dL) for a cantilever of length L bent by F at its end; the
    checker compares it with E pi r^4 / 4 (:47, :82)."""
    return F * L ** 3 / (3 * dL)


def surface_energy(measurements: dict) -> float:
    """tests/physics/cubes_lj/checkE.py:67-77: the last SurfSurfEnergy of a measurement file
    (J); the checker wants it within 1 % of -1.7870e-19 J for its two cubes."""
    return float(measurements["SurfSurfEnergy"][-1])
