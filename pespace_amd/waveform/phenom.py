"""IMRPhenomD / PhenomHM waveform models for the hip backend.

Same interface shape as the reference's waveform adaptors
(pespace/waveform/interface_ripple.py:27-90): constructed on the frequency
samples with a reference frequency, ``required_params``, ``get_waveform(params)``
returning ``{"plus","cross","tf"}`` (or a mode-keyed dict of those) with t_c = 0,
and the taichi-style ``update_waveform(dict)`` / ``waveform_container`` pair used
by the bilby interface (interface_bilby.py:97-101).  The model itself runs in
the HIP library: amplitude and phase on ``n_nodes`` log-spaced nodes (or on
every bin when ``n_nodes == 0``), cubic-Hermite interpolated inside the fused
response kernel.
"""

from __future__ import annotations

import numpy as np

from ..core import MODE_IDS, PARAM_NAMES, LikelihoodCore

HM_MODES = ("22", "21", "33", "32", "44", "43")


def _mass_params(params):
    """chirp_mass / mass_ratio (m2/m1 <= 1) from any of the reference's mass conventions."""
    if "chirp_mass" in params and "mass_ratio" in params:
        return float(params["chirp_mass"]), float(params["mass_ratio"])
    if "mass_1" in params and "mass_2" in params:  # interface_ripple.py:188-198
        m1, m2 = float(params["mass_1"]), float(params["mass_2"])
        m1, m2 = max(m1, m2), min(m1, m2)
        return (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2, m2 / m1
    if "chirp_mass" in params and "symmetric_mass_ratio" in params:  # interface_ripple.py:106-116
        eta = float(params["symmetric_mass_ratio"])
        eta = min(eta, 0.25)
        d = np.sqrt(1.0 - 4.0 * eta)
        return float(params["chirp_mass"]), (1.0 - d) / (1.0 + d)
    raise KeyError("need chirp_mass+mass_ratio, mass_1+mass_2 or chirp_mass+symmetric_mass_ratio")


class PhenomWaveformBase:
    modes = ("22",)

    def __init__(self, frequencies, reference_frequency=0.0, n_nodes=1024, f_node_min=None, f_node_max=None):
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.reference_frequency = float(reference_frequency)
        self.n_nodes = int(n_nodes)
        self.f_node_min = float(self.frequencies.min()) if f_node_min is None else float(f_node_min)
        self.f_node_max = float(self.frequencies.max()) if f_node_max is None else float(f_node_max)
        self._params = None
        self._core = None

    @property
    def required_params(self):
        return ("chirp_mass", "mass_ratio", "chi_1", "chi_2", "luminosity_distance", "reference_phase",
                "inclination")

    # taichi-style interface (interface_bilby.py:97-101)
    def update_waveform(self, params):
        mc, q = _mass_params(params)
        self._params = dict(chirp_mass=mc, mass_ratio=q, chi_1=float(params["chi_1"]),
                            chi_2=float(params["chi_2"]),
                            luminosity_distance=float(params["luminosity_distance"]),
                            reference_phase=float(params["reference_phase"]),
                            inclination=float(params["inclination"]))

    @property
    def waveform_container(self):
        if self._params is None:
            raise RuntimeError("call update_waveform(params) first")
        return self

    def parameter_row(self, lam, beta, psi, tc):
        p = dict(self._params)
        p.update(ecliptic_longitude=float(lam), ecliptic_latitude=float(beta), polarization=float(psi),
                 coalescence_time=float(tc))
        return np.array([p[k] for k in PARAM_NAMES], dtype=np.float64)

    def _grid_core(self):
        if self._core is None:
            core = LikelihoodCore(modes=self.modes, reference_frequency=self.reference_frequency,
                                  n_nodes=self.n_nodes, f_node_min=self.f_node_min, f_node_max=self.f_node_max)
            n = self.frequencies.size
            core.add_detector(0, 2.5e9, 0.0, 0.0, 15, True, self.frequencies, np.zeros((3, n), np.complex128),
                              np.ones((3, n)), [0, 0, 0], df=1.0)
            self._core = core
        return self._core

    def get_waveform(self, params):
        """Container on ``frequencies`` (torch tensors on the GPU), t_c = 0."""
        p = params if isinstance(params, dict) else params._asdict()
        self.update_waveform(p)
        out = self._grid_core().waveform(0, self.parameter_row(0.0, 0.0, 0.0, 0.0)[None, :])[0]
        res = {}
        for i, m in enumerate(self.modes):
            res[m] = {"plus": out[i, :, 0] + 1j * out[i, :, 1], "cross": out[i, :, 2] + 1j * out[i, :, 3],
                      "tf": out[i, :, 4]}
        return res[self.modes[0]] if len(self.modes) == 1 else res

    def __call__(self, params):
        return self.get_waveform(params)


class IMRPhenomD(PhenomWaveformBase):
    """(2,2) mode, Khan et al. 2016 / Husa et al. 2016."""

    modes = ("22",)


class IMRPhenomHM(PhenomWaveformBase):
    """PhenomD plus the (2,1),(3,3),(3,2),(4,4),(4,3) modes (London et al. 2018 construction)."""

    def __init__(self, frequencies, reference_frequency=0.0, modes=HM_MODES, n_nodes=1024, f_node_min=None,
                 f_node_max=None):
        super().__init__(frequencies, reference_frequency, n_nodes, f_node_min, f_node_max)
        modes = tuple(modes)
        if "22" not in modes:
            modes = ("22",) + modes
        for m in modes:
            if m not in MODE_IDS:
                raise ValueError(f"unsupported mode {m}")
        self.modes = modes
