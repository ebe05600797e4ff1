"""The SMC drivers' known answers on the GPU: a Gaussian likelihood under a uniform box prior
has log Z = -log(box volume) (the Gaussian's mass inside the box is 1 to 1e-300) and the
Gaussian itself as posterior.  The same checks as the CPU tests, with every tensor and the
random generators on the GPU (IRMH and NUTS mutations)."""

import math
from collections import namedtuple

import numpy as np
import pytest
import torch

from pespace_amd.inference import nuts, smc

pytestmark = pytest.mark.gpu

MEAN, COV = [1.0, -2.0], [[0.04, 0.012], [0.012, 0.09]]
PRIOR = {"u": {"distribution": "Uniform", "low": -10.0, "high": 10.0},
         "v": {"distribution": "Uniform", "low": -10.0, "high": 10.0}}


class _GaussLikeCuda:
    """Normalised Gaussian density in two parameters, batched, value and gradient on the GPU."""

    likelihood_param_type = namedtuple("P", ("u", "v"))

    def __init__(self):
        self.mean = torch.tensor(MEAN, dtype=torch.float64, device="cuda")
        self.cov = torch.tensor(COV, dtype=torch.float64, device="cuda")
        self.prec = torch.linalg.inv(self.cov.cpu()).to("cuda")
        self.chol = torch.linalg.cholesky(self.cov.cpu()).to("cuda")

    def log_likelihood_batch(self, rows):
        return smc.gaussian_logpdf_chol(rows, self.mean, self.chol)

    def log_likelihood_and_grad(self, rows):
        d = rows - self.mean
        return self.log_likelihood_batch(rows), -(d @ self.prec)


def _check(out, n_evals=None, tol_z=0.1, tol_mean=0.02, rtol_cov=0.1):
    betas = [lv["beta"] for lv in out["levels"]]
    assert betas[-1] == 1.0 and np.all(np.diff(betas) > 0)
    assert abs(out["log_evidence"] + math.log(400.0)) < tol_z, out["log_evidence"]
    post = np.stack([out["posteriors"]["u"].cpu().numpy(), out["posteriors"]["v"].cpu().numpy()], 1)
    np.testing.assert_allclose(post.mean(0), MEAN, atol=tol_mean)
    np.testing.assert_allclose(np.cov(post.T), COV, rtol=rtol_cov, atol=3e-3)


def test_irmh_smc_known_answer_on_gpu():
    s = smc.TemperedSMC(_GaussLikeCuda(), PRIOR, num_particles=4000, target_ess=0.9, num_mcmc_steps=20, rkey_seed=3,
                        device="cuda")
    out = s.run()
    assert out["posteriors"]["u"].device.type == "cuda"
    _check(out)


def test_nuts_smc_known_answer_on_gpu():
    s = nuts.TemperedSMCNUTS(_GaussLikeCuda(), PRIOR, num_particles=2000, target_ess=0.9, num_mcmc_steps=2,
                             step_size=0.5, max_num_doublings=6, rkey_seed=7, device="cuda")
    out = s.run()
    _check(out, tol_z=0.15, tol_mean=0.03, rtol_cov=0.15)
    assert all(lv["acceptance"] > 0.3 for lv in out["levels"])
