"""Analytic Keplerian constellation orbits (parameters only on the host).

Same constructors as pespace/detector/orbit.py:67-241; the orbit itself is
evaluated inside the HIP kernels at t = tf + t_c for every (bin, mode, row).
"""

from __future__ import annotations

from ..utils.constants import C_SI, DAY_SI, PI

BACKENDS = ("hip",)


class OrbitModelBase:
    kind = None


class KaplerianHeliocentric(OrbitModelBase):
    """LISA / Taiji heliocentric orbit (orbit.py:155-241)."""

    kind = 0

    def __init__(self, arm_length, rotation_initial=0.0, revolution_initial=0.0, backend="hip"):
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        self.arm_length = float(arm_length)
        self.rotation_initial = float(rotation_initial)
        self.revolution_initial = float(revolution_initial)
        self.backend = backend

    def arm_length_sec(self, time=0.0):
        return self.arm_length / C_SI


class KeplerianGeocentric(OrbitModelBase):
    """TianQin geocentric orbit (orbit.py:67-152); the device uses the TianQin rotation and reference source."""

    kind = 1

    def __init__(self, arm_length, rotation_initial=0.0, revolution_initial=0.0,
                 omega_rotation=2 * PI / (3.65 * DAY_SI), lambda_ref=120.5 / 180 * PI, beta_ref=-4.7 / 180 * PI,
                 backend="hip"):
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        default = (2 * PI / (3.65 * DAY_SI), 120.5 / 180 * PI, -4.7 / 180 * PI)
        if (omega_rotation, lambda_ref, beta_ref) != default:
            raise ValueError("the hip backend implements the TianQin rotation rate and reference source only")
        self.arm_length = float(arm_length)
        self.rotation_initial = float(rotation_initial)
        self.revolution_initial = float(revolution_initial)
        self.backend = backend

    def arm_length_sec(self, time=0.0):
        return self.arm_length / C_SI


def available_orbit_models(backend="hip"):
    """orbit.py:442-446."""
    return {
        "LISA_analytic": KaplerianHeliocentric(2.5e9, 0.0, -PI / 9, backend=backend),
        "Taiji_analytic": KaplerianHeliocentric(3.0e9, 0.0, PI / 9, backend=backend),
        "Tianqin_analytic": KeplerianGeocentric(1.7e8, 0.0, 0.0, backend=backend),
    }
