"""Analytic TDI noise PSDs (setup-time, host numpy).

Same models and call signature as pespace/detector/noise.py:12-284:
``model(frequencies, tdi_channels, scaling, tdi_generation) -> {channel: S(f)}``.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass

import numpy as np

from ..utils.constants import C_SI, MRSUN_SI, MTSUN_SI, PC_SI, PI


class FrequencyDomainNoiseModel(ABC):
    """noise.py:12-46."""

    @abstractmethod
    def power_spectral_density_array(self, frequencies, tdi_channels, scaling, tdi_generation):
        ...

    def __call__(self, frequencies, tdi_channels, scaling, tdi_generation):
        return self.power_spectral_density_array(frequencies, tdi_channels, scaling, tdi_generation)


@dataclass
class AnalyticPowerSpectralDensity(FrequencyDomainNoiseModel):
    """OMS + acceleration noise in TDI X/Y/Z and A/E/T (noise.py:49-207, arXiv:2108.01167)."""

    OMS_noise_level: float
    acc_noise_level: float
    arm_length_sec: float

    def _S_oms(self, f):
        return self.OMS_noise_level * (1.0 + (2.0e-3 / f) ** 4) * (2.0 * PI * f / C_SI) ** 2

    def _S_acc(self, f):
        return self.acc_noise_level * (1.0 + (0.4e-3 / f) ** 2) * (1.0 + (f / 8e-3) ** 4) / (2 * PI * f * C_SI) ** 2

    def power_spectral_density_array(self, frequencies, tdi_channels, scaling, tdi_generation):
        f = np.asarray(frequencies, dtype=np.float64)
        S_oms = self._S_oms(f)
        S_acc = self._S_acc(f)
        if scaling:
            fac = (PC_SI / (MRSUN_SI * MTSUN_SI)) ** 2
            S_oms = S_oms * fac
            S_acc = S_acc * fac
        L = self.arm_length_sec
        if tdi_generation == "1.5":
            pre = 1.0
        elif tdi_generation == "2.0":
            pre = 4.0 * np.sin(4 * PI * f * L) ** 2
        else:
            raise ValueError("The TDI generation {} is unknown".format(tdi_generation))
        s2 = np.sin(2 * PI * f * L) ** 2
        out = {}
        for chan in tdi_channels:
            if chan in ("X", "Y", "Z"):
                psd = 16 * s2 * (S_oms + (3 + np.cos(4 * PI * f * L)) * S_acc)
            elif chan in ("A", "E"):
                psd = 8 * s2 * ((2 + np.cos(2 * PI * f * L)) * S_oms
                                + (6 + 4 * np.cos(2 * PI * f * L) + 2 * np.cos(4 * PI * f * L)) * S_acc)
            elif chan == "T":
                sp = np.sin(PI * f * L) ** 2
                psd = 32 * s2 * sp * (S_oms + 4 * sp * S_acc)
            else:
                raise ValueError(f"Unknown TDI channel {chan}.")
            out[chan] = psd * pre
        return out


class TianQinAnalyticPowerSpectralDensity(AnalyticPowerSpectralDensity):
    """noise.py:210-265 (arXiv:2309.15020)."""

    def _S_oms(self, f):
        return self.OMS_noise_level * (2.0 * PI * f / C_SI) ** 2

    def _S_acc(self, f):
        return self.acc_noise_level * (1.0 + 0.1e-3 / f) / (2 * PI * f * C_SI) ** 2


available_noise_models = {
    "LISA_SciRDv1": AnalyticPowerSpectralDensity((15.0e-12) ** 2, (3.0e-15) ** 2, 2.5e9 / C_SI),
    "Taiji_TDC": AnalyticPowerSpectralDensity((8.0e-12) ** 2, (3.0e-15) ** 2, 3.0e9 / C_SI),
    "Tianqin_GWSpace": TianQinAnalyticPowerSpectralDensity((1.0e-12) ** 2, (1.0e-15) ** 2, 1.7e8 / C_SI),
}
