from .likelihood import (LikelihoodBilbyInterfaceHip, WhittleLikelihoodHip, WhittleLnL, compute_whittle_likelihood,
                         noise_weighted_inner_product)
from .smc import Cosine, Prior, Sine, TemperedSMC

__all__ = ["WhittleLikelihoodHip", "WhittleLnL", "LikelihoodBilbyInterfaceHip", "compute_whittle_likelihood",
           "noise_weighted_inner_product", "TemperedSMC", "Prior", "Sine", "Cosine"]
from .fisher import fisher_matrix, sample_from_fisher, wrap_angle_params

__all__ += ["fisher_matrix", "sample_from_fisher", "wrap_angle_params"]
from .nuts import TemperedSMCNUTS, nuts_step, softabs_inverse_metric

__all__ += ["TemperedSMCNUTS", "nuts_step", "softabs_inverse_metric"]
