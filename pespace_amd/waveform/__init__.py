from .phenom import IMRPhenomD, IMRPhenomHM, PhenomWaveformBase

__all__ = ["IMRPhenomD", "IMRPhenomHM", "PhenomWaveformBase"]
