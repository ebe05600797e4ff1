"""Adaptive-tempered SMC driver (pespace_amd.inference.smc) on the CPU.

The primitives are pinned against closed forms and brute force: log ESS of equal and one-hot
weights; the tempering increment against a dense scan of the ESS curve; systematic
resampling against its defining property (particle i is copied floor(N W_i) or
ceil(N W_i) times, and the copies follow the cumulative weights); the Gaussian fit against
numpy's covariance and scipy's log-density.  End to end, a Gaussian likelihood under a wide
uniform prior has a known evidence (log Z = -log of the box volume) and posterior (the
Gaussian itself).
"""

from collections import namedtuple

import math

import numpy as np
import pytest
import torch
from scipy import stats

from pespace_amd.inference import smc


def test_log_ess_closed_forms():
    n = 1000
    lw = torch.full((n,), -3.7, dtype=torch.float64)
    assert abs(float(smc.log_ess(lw)) - math.log(n)) < 1e-12
    one = torch.full((n,), -math.inf, dtype=torch.float64)
    one[17] = 2.0
    assert abs(float(smc.log_ess(one))) < 1e-12
    rng = np.random.default_rng(0)
    w = torch.as_tensor(rng.normal(size=n))
    ess = float(torch.exp(smc.log_ess(w)))
    W = np.exp(w.numpy())
    assert abs(ess - W.sum() ** 2 / (W ** 2).sum()) < 1e-9 * ess


@pytest.mark.parametrize("beta", [0.0, 0.3, 0.97])
def test_tempering_increment_hits_target(beta):
    rng = np.random.default_rng(1)
    logl = torch.as_tensor(-0.5 * rng.chisquare(11, 4096) * 40.0)
    logl[5] = -math.inf  # a particle outside the likelihood's domain
    target = 0.9
    d = float(smc.tempering_increment(logl, torch.tensor(beta, dtype=torch.float64), target))
    assert 0.0 <= d <= 1.0 - beta + 1e-15

    def f(x):
        return float(smc.log_ess(torch.nan_to_num(x * logl))) - math.log(4096 * target)

    if f(1.0 - beta) >= 0.0:
        assert d == pytest.approx(1.0 - beta, abs=1e-15)
    else:
        assert abs(f(d)) <= 1e-4
        grid = np.linspace(0.0, 1.0 - beta, 20001)
        vals = np.array([f(g) for g in grid])
        root = grid[np.argmax(vals < 0.0)]  # first grid point past the (unique) root
        assert abs(d - root) <= 2 * (grid[1] - grid[0])


def test_tempering_takes_the_rest_when_ess_allows():
    logl = torch.full((256,), -1.0, dtype=torch.float64)  # equal weights at any delta
    d = float(smc.tempering_increment(logl, torch.tensor(0.4, dtype=torch.float64), 0.9))
    assert d == pytest.approx(0.6)


def test_systematic_resampling_counts():
    rng = np.random.default_rng(2)
    n = 997
    lw = torch.as_tensor(rng.normal(size=n) * 3.0)
    W = torch.softmax(lw, 0).numpy()
    for u in (0.0, 0.25, 0.999):
        idx = smc.systematic_resample(lw, torch.tensor(u, dtype=torch.float64)).numpy()
        counts = np.bincount(idx, minlength=n)
        assert counts.sum() == n
        assert np.all(np.abs(counts - n * W) < 1.0 + 1e-9)
        assert np.all(np.diff(idx) >= 0)
        # restatement: index of the first cumulative weight >= (i + u) / N
        cs = np.cumsum(W)
        ref = np.minimum(np.searchsorted(cs, (np.arange(n) + u) / n), n - 1)
        # cumulative sums in torch and numpy may differ in the last bit near a boundary
        assert np.mean(idx == ref) > 0.995


def test_fit_gaussian_and_logpdf():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(4, 4))
    x = rng.normal(size=(500, 4)) @ A.T + np.array([1.0, -2.0, 0.5, 3.0])
    mean, cov, chol = smc.fit_gaussian(torch.as_tensor(x), 1e-6)
    np.testing.assert_allclose(mean.numpy(), x.mean(0), rtol=1e-12)
    np.testing.assert_allclose(cov.numpy(), np.cov(x.T) + 1e-6 * np.eye(4), rtol=1e-12)
    np.testing.assert_allclose((chol @ chol.T).numpy(), cov.numpy(), rtol=1e-12, atol=1e-14)
    assert np.allclose(np.triu(chol.numpy(), 1), 0.0)
    z = rng.normal(size=(7, 4))
    np.testing.assert_allclose(smc.gaussian_from_normals(mean, chol, torch.as_tensor(z)).numpy(),
                               mean.numpy() + z @ chol.numpy().T, rtol=1e-12)
    lp = smc.gaussian_logpdf_chol(torch.as_tensor(x[:20]), mean, chol).numpy()
    ref = stats.multivariate_normal(mean.numpy(), cov.numpy()).logpdf(x[:20])
    np.testing.assert_allclose(lp, ref, rtol=1e-10)


def test_prior_validation_messages():
    names = ("a", "b", "c")
    ok = {"a": {"distribution": "Uniform", "low": 0.0, "high": 1.0}, "b": {"fixed": 2.0},
          "c": {"transform": lambda a, b: a * b, "from": ("a", "b")}}
    p = smc.Prior(ok, names)
    assert p.sampled_names == ("a",)
    x = torch.tensor([[0.25], [0.5]], dtype=torch.float64)
    np.testing.assert_allclose(p.rows(x).numpy(), [[0.25, 2.0, 0.5], [0.5, 2.0, 1.0]])
    with pytest.raises(ValueError, match="not declared before them"):
        smc.Prior({"c": {"transform": lambda a: a, "from": ("a",)}, "a": {"fixed": 1.0}, "b": {"fixed": 1.0}},
                  names)
    with pytest.raises(ValueError, match="declared more than once"):
        smc.Prior({"a": {"fixed": 1.0}, "b": {"fixed": 1.0}, "c": {"fixed": 1.0},
                   "x": {"transform": lambda a: a, "from": ("a",)},
                   ("a",): {"transform": lambda b: b, "from": ("b",)}}, names)
    with pytest.raises(ValueError, match="Missing"):
        smc.Prior({"a": {"fixed": 1.0}}, names)
    with pytest.raises(ValueError, match="must contain"):
        smc.Prior({"a": {"low": 1.0}, "b": {"fixed": 1.0}, "c": {"fixed": 1.0}}, names)


@pytest.mark.parametrize("spec", [
    {"distribution": "Uniform", "low": -1.0, "high": 3.0},
    {"distribution": "LogUniform", "low": 0.1, "high": 10.0},
    {"distribution": "Sine"},
    {"distribution": smc.Cosine},
    {"distribution": "Normal", "loc": 1.0, "scale": 0.5},
])
def test_prior_densities_normalised_and_sampling_matches(spec):
    p = smc.Prior({"a": spec}, ("a",))
    d = p.sampled["a"]
    lo, hi = {"Uniform": (-1, 3), "LogUniform": (0.1, 10), "Sine": (0, math.pi),
              "Cosine": (-math.pi / 2, math.pi / 2), "Normal": (-4, 6)}[d.kind]
    g = torch.linspace(lo, hi, 200001, dtype=torch.float64)
    dens = torch.exp(d.log_prob(g)).numpy()
    assert abs(np.trapezoid(dens, g.numpy()) - 1.0) < 1e-6
    gen = torch.Generator().manual_seed(5)
    s = p.sample(200000, gen, "cpu")[:, 0].numpy()
    cdf = np.cumsum(dens) * (g[1] - g[0]).item()
    ks = np.max(np.abs(np.searchsorted(np.sort(s), g.numpy()) / s.size - cdf))
    assert ks < 0.01
    assert torch.all(torch.isfinite(d.log_prob(torch.as_tensor(s[:1000]))))


class _GaussLike:
    """Normalised Gaussian density in two parameters, batched like WhittleLikelihoodHip."""

    likelihood_param_type = namedtuple("P", ("u", "v"))

    def __init__(self, mean, cov):
        self.mean = torch.as_tensor(mean, dtype=torch.float64)
        self.cov = torch.as_tensor(cov, dtype=torch.float64)
        self.chol = torch.linalg.cholesky(self.cov)

    def log_likelihood_batch(self, rows):
        return smc.gaussian_logpdf_chol(rows, self.mean, self.chol)


def test_smc_evidence_and_posterior_of_a_gaussian():
    mean, cov = [1.0, -2.0], [[0.04, 0.012], [0.012, 0.09]]
    lk = _GaussLike(mean, cov)
    prior = {"u": {"distribution": "Uniform", "low": -10.0, "high": 10.0},
             "v": {"distribution": "Uniform", "low": -10.0, "high": 10.0}}
    sampler = smc.TemperedSMC(lk, prior, num_particles=4000, target_ess=0.9, num_mcmc_steps=20, rkey_seed=3,
                              device="cpu")
    out = sampler.run()
    betas = [lv["beta"] for lv in out["levels"]]
    assert betas[-1] == 1.0 and np.all(np.diff(betas) > 0)
    # Z = (1/400) * (Gaussian mass inside the box, 1 to 1e-300)
    assert abs(out["log_evidence"] + math.log(400.0)) < 0.1
    post = np.stack([out["posteriors"]["u"].numpy(), out["posteriors"]["v"].numpy()], 1)
    np.testing.assert_allclose(post.mean(0), mean, atol=0.02)
    np.testing.assert_allclose(np.cov(post.T), cov, rtol=0.1, atol=2e-3)
    assert all(0.0 < lv["acceptance"] <= 1.0 for lv in out["levels"])
    assert out["n_likelihood_evaluations"] == 4000 * (1 + 20 * len(betas))
