"""FFEA meso unit system and the simulation parameters the massless step reads.

Values follow include/dimensions.h:30-41 (mesoDimensions) and the conversions of
SimulationParams::assign (src/SimulationParams.cpp:222-230 dt and epsilon,
:336-339 kT, :402-405 stokes_visc) and of World (src/World.cpp:2565-2567, blob scale).
"""
from __future__ import annotations

import dataclasses
import math

MESO_LENGTH = 1.7e-10
MESO_ENERGY = 4.141945559999999e-21
MESO_MASS = 1.994307387553024e-26
MESO_VOLUME = 4.913e-30
MESO_TIME = 3.7302915560886126e-13
MESO_PRESSURE = 8.4305832688784826e8
MESO_FORCE = 2.4364705882352941e-11
MESO_AREA = 2.89e-20
MESO_VELOCITY = 4.5572845297447225e2

KB = 1.380649e-23
PI = 3.141592653589793


@dataclasses.dataclass
class SimParams:
    """The SimulationParams fields on the CG_nomass path, in meso units.

    Construct with :meth:`from_si` to apply the reference's unit conversions.
    """

    dt: float
    kT: float
    epsilon2: float
    max_iterations_cg: int = 1000
    calc_noise: int = 1
    calc_stokes: int = 1
    stokes_visc: float = 1e-3 / (MESO_PRESSURE * MESO_TIME)
    rng_seed: int = 44

    @classmethod
    def from_si(cls, dt=1e-14, kT=KB * 300.0, epsilon=1e-6, max_iterations_cg=1000,
                calc_noise=1, calc_stokes=1, stokes_visc=1e-3, rng_seed=44):
        return cls(dt=dt / MESO_TIME, kT=kT / MESO_ENERGY, epsilon2=epsilon * epsilon,
                   max_iterations_cg=int(max_iterations_cg), calc_noise=int(calc_noise),
                   calc_stokes=int(calc_stokes),
                   stokes_visc=stokes_visc / (MESO_PRESSURE * MESO_TIME), rng_seed=int(rng_seed))


def blob_scale(scale_si: float) -> float:
    """`scale /= mesoDimensions::length` (src/World.cpp:2565-2567)."""
    s = scale_si
    s /= MESO_LENGTH
    return s


def ns_per_day(steps_per_s: float, dt_si: float = 1e-14) -> float:
    return steps_per_s * dt_si * 1e9 * 86400.0


def stokes_drag(radius_meso: float, stokes_visc_meso: float) -> float:
    """6 pi eta r (src/Blob.cpp:408-412)."""
    return 6.0 * math.pi * stokes_visc_meso * radius_meso
