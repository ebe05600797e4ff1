"""Analytic Keplerian constellation orbits, hip backend.

Same constructors and methods as pespace/detector/orbit.py:67-241:
``get_constellation_vectors(time)`` returns the link unit vectors n1, n2, n3 and the node
positions x1, x2, x3 (SSB frame, seconds) at an array of times, computed on the GPU by the
same device function the response kernels use; ``arm_length_sec(time)`` is the constant arm
length in seconds.  Inside the fused likelihood kernel the orbit is evaluated at
t = tf + t_c for every (bin, mode, row) from the constants in ``orbit_params``.
"""

from __future__ import annotations

from collections import namedtuple

import numpy as np

from ..utils.constants import C_SI, DAY_SI, PI, YEAR_SI

BACKENDS = ("hip",)

# orbit.py:17-24 (taichi struct), _jax_impl/_orbit_jax.py:16-18 (JAX tuple)
ConstellationVectorTuple = namedtuple("ConstellationVectorTuple", ["n1", "n2", "n3", "x1", "x2", "x3"])


class OrbitModelBase:
    kind = None

    def _vectors(self, time, **geo):
        from ..core import orbit_vectors
        v = orbit_vectors(self.kind, self.arm_length, self.rotation_initial, self.revolution_initial, time, **geo)
        return ConstellationVectorTuple(*(v[..., i, :] for i in range(6)))

    def arm_length_sec(self, time=0.0):
        """Constant arm length in seconds (orbit.py:323-342); ``time`` is kept for the interface."""
        return self.orbit_params.static_arm_sec


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
        arm_sec = self.arm_length / C_SI
        kap = [self.rotation_initial + i * 2 * PI / 3 for i in range(3)]
        OrbitParams = namedtuple("OrbitParams", ["static_arm_sec", "rotation_initial", "revolution_initial",
                                                 "r_prime", "omega", "ck1", "sk1", "ck2", "sk2", "ck3", "sk3"])
        self.orbit_params = OrbitParams(
            static_arm_sec=arm_sec, rotation_initial=self.rotation_initial,
            revolution_initial=self.revolution_initial, r_prime=arm_sec / (2 * np.sqrt(3)), omega=2 * PI / YEAR_SI,
            ck1=np.cos(kap[0]), sk1=np.sin(kap[0]), ck2=np.cos(kap[1]), sk2=np.sin(kap[1]), ck3=np.cos(kap[2]),
            sk3=np.sin(kap[2]))

    def get_constellation_vectors(self, time):
        """n1..n3, x1..x3 at ``time`` (array-like), each a (..., 3) tensor (orbit.py:363-419)."""
        return self._vectors(time)


class KeplerianGeocentric(OrbitModelBase):
    """TianQin-like geocentric orbit (orbit.py:67-152); any rotation rate and reference source."""

    kind = 1

    def __init__(self, arm_length, rotation_initial=0.0, revolution_initial=0.0,
                 omega_rotation=2 * PI / (3.65 * DAY_SI), lambda_ref=120.5 / 180 * PI, beta_ref=-4.7 / 180 * PI,
                 backend="hip"):
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        self.arm_length = float(arm_length)
        self.rotation_initial = float(rotation_initial)
        self.revolution_initial = float(revolution_initial)
        self.omega_rotation = float(omega_rotation)
        self.lambda_ref = float(lambda_ref)
        self.beta_ref = float(beta_ref)
        self.backend = backend
        OrbitParams = namedtuple("OrbitParams", ["static_arm_sec", "r_det_sec", "rotation_initial",
                                                 "revolution_initial", "omega_rotation", "omega_revolution",
                                                 "slam_ref", "clam_ref", "sbeta_ref", "cbeta_ref"])
        arm_sec = self.arm_length / C_SI
        self.orbit_params = OrbitParams(
            static_arm_sec=arm_sec, r_det_sec=arm_sec / np.sqrt(3), rotation_initial=self.rotation_initial,
            revolution_initial=self.revolution_initial, omega_rotation=self.omega_rotation,
            omega_revolution=2 * PI / YEAR_SI, slam_ref=np.sin(self.lambda_ref), clam_ref=np.cos(self.lambda_ref),
            sbeta_ref=np.sin(self.beta_ref), cbeta_ref=np.cos(self.beta_ref))

    @property
    def geo_params(self):
        """Rotation rate and reference source, as keyword arguments of the C-ABI wrappers."""
        return dict(omega_rotation=self.omega_rotation, lambda_ref=self.lambda_ref, beta_ref=self.beta_ref)

    def get_constellation_vectors(self, time):
        """n1..n3, x1..x3 at ``time`` (array-like), each a (..., 3) tensor (orbit.py:260-322)."""
        return self._vectors(time, **self.geo_params)


def available_orbit_models(backend="hip"):
    """orbit.py:442-446."""
    return {
        "LISA_analytic": KaplerianHeliocentric(2.5e9, 0.0, -PI / 9, backend=backend),
        "Taiji_analytic": KaplerianHeliocentric(3.0e9, 0.0, PI / 9, backend=backend),
        "Tianqin_analytic": KeplerianGeocentric(1.7e8, 0.0, 0.0, backend=backend),
    }
