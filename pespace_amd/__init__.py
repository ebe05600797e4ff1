"""pespace_amd: MI355X-native FD lnL(+grad) hot path for space-borne GW detectors.

Host API mirrors nnrui/pespace (``backend="hip"``); compute runs in the gfx950
HIP library ``libpespace_amd.so`` through the C-ABI of include/pespace_amd.h.
"""

from .core import MODE_IDS, N_PARAMS, PARAM_NAMES, LikelihoodCore

__all__ = ["LikelihoodCore", "PARAM_NAMES", "N_PARAMS", "MODE_IDS"]
