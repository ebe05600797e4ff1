from .likelihood import (LikelihoodBilbyInterfaceHip, WhittleLikelihoodHip, WhittleLnL, compute_whittle_likelihood,
                         noise_weighted_inner_product)
from .smc import Cosine, Prior, Sine, TemperedSMC

__all__ = ["WhittleLikelihoodHip", "WhittleLnL", "LikelihoodBilbyInterfaceHip", "compute_whittle_likelihood",
           "noise_weighted_inner_product", "TemperedSMC", "Prior", "Sine", "Cosine"]
