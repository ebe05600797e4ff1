from .likelihood import (LikelihoodBilbyInterfaceHip, WhittleLikelihoodHip, WhittleLnL, compute_whittle_likelihood,
                         noise_weighted_inner_product)

__all__ = ["WhittleLikelihoodHip", "WhittleLnL", "LikelihoodBilbyInterfaceHip", "compute_whittle_likelihood",
           "noise_weighted_inner_product"]
