"""NUTS mutation and its SMC driver (pespace_amd.inference.nuts) on the CPU.

Pins: the unconstrained map round-trips and its log-Jacobian equals a numerical derivative;
SoftAbs reduces to |lambda| for large |alpha lambda| and to 1/alpha near zero; one batched
NUTS transition leaves a correlated Gaussian invariant (exact samples stay distributed as the
target) and mixes a point mass into it; end to end, SMC with NUTS mutations recovers the
evidence and posterior of a Gaussian likelihood under a box prior, as the IRMH driver does.
"""

from collections import namedtuple

import math

import numpy as np
import pytest
import torch

from pespace_amd.inference import nuts, smc


def test_unconstrained_round_trip_and_log_det():
    prior = smc.Prior({"a": {"distribution": "Uniform", "low": -2.0, "high": 5.0}, "b": {"distribution": "Sine"},
                       "c": {"distribution": "Normal", "loc": 0.0, "scale": 1.0}}, ("a", "b", "c"))
    u = nuts.Unconstrained(prior)
    x = prior.sample(1000, torch.Generator().manual_seed(1), "cpu")
    z = u.to_z(x)
    np.testing.assert_allclose(u.to_x(z).numpy(), x.numpy(), rtol=1e-12, atol=1e-12)
    h = 1e-6
    zc = z[:5].clone()
    num = []
    for i in range(3):
        zp, zm = zc.clone(), zc.clone()
        zp[:, i] += h
        zm[:, i] -= h
        num.append(((u.to_x(zp) - u.to_x(zm))[:, i] / (2 * h)).abs())
    np.testing.assert_allclose(u.log_det(zc).numpy(), torch.log(torch.stack(num, 1)).sum(1).numpy(), atol=1e-8)


def test_softabs_limits():
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)))
    lam = np.array([1e4, -3e3, 1e-12, 50.0])
    H = torch.as_tensor((Q * lam) @ Q.T)[None]
    A = nuts.softabs_inverse_metric(H, 0.3)[0].numpy()
    soft = np.where(np.abs(lam) * 0.3 < 1e-8, 1 / 0.3, lam / np.tanh(0.3 * lam))
    np.testing.assert_allclose(A, (Q / soft) @ Q.T, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(soft[:2], np.abs(lam[:2]), rtol=1e-12)
    assert abs(soft[2] - 1 / 0.3) < 1e-9
    assert np.all(np.linalg.eigvalsh(A) > 0)


def _gauss(mean, cov):
    mean = torch.as_tensor(mean, dtype=torch.float64)
    prec = torch.linalg.inv(torch.as_tensor(cov, dtype=torch.float64))

    def f(z):
        r = z - mean
        g = -(r @ prec)
        return 0.5 * (r * g).sum(1), g
    return f


@pytest.mark.parametrize("start", ["exact", "point"])
def test_nuts_keeps_and_reaches_a_gaussian(start):
    mean, cov = np.array([1.0, -2.0, 0.5]), np.array([[1.0, 0.8, 0.1], [0.8, 1.0, 0.0], [0.1, 0.0, 0.25]])
    f = _gauss(mean, cov)
    n = 20000
    g = torch.Generator().manual_seed(4)
    if start == "exact":
        z = torch.as_tensor(np.random.default_rng(5).multivariate_normal(mean, cov, n))
    else:
        z = torch.as_tensor(np.tile(mean + 1.0, (n, 1)))
    A = torch.as_tensor(cov)[None].expand(n, 3, 3).clone()
    lp, gr = f(z)
    # from a point the potential energy relaxes through the momentum refreshes (~1 per step)
    steps = 1 if start == "exact" else 10
    for _ in range(steps):
        z, lp, gr, info = nuts.nuts_step(f, z, lp, gr, A, 0.4, g, max_num_doublings=6, divergence_threshold=1000.0)
        assert not bool(info["is_divergent"].any())
        assert float(info["acceptance_rate"].mean()) > 0.7
    zz = z.numpy()
    np.testing.assert_allclose(zz.mean(0), mean, atol=0.04)
    np.testing.assert_allclose(np.cov(zz.T), cov, atol=0.05)
    lp_ref, _ = f(z)
    np.testing.assert_allclose(lp.numpy(), lp_ref.numpy(), rtol=1e-12)


class _GaussLike:
    likelihood_param_type = namedtuple("P", ("u", "v"))

    def __init__(self, mean, cov):
        self.f = _gauss(mean, cov)
        self.norm = -0.5 * math.log(np.linalg.det(2 * np.pi * np.asarray(cov)))

    def log_likelihood_and_grad(self, rows):
        v, g = self.f(rows)
        return v + self.norm, g


def test_smc_with_nuts_evidence_and_posterior():
    mean, cov = [1.0, -2.0], [[0.04, 0.012], [0.012, 0.09]]
    prior = {"u": {"distribution": "Uniform", "low": -10.0, "high": 10.0},
             "v": {"distribution": "Uniform", "low": -10.0, "high": 10.0}}
    s = nuts.TemperedSMCNUTS(_GaussLike(mean, cov), prior, num_particles=2000, target_ess=0.9, num_mcmc_steps=2,
                             step_size=0.5, max_num_doublings=6, rkey_seed=7, device="cpu")
    out = s.run()
    betas = [lv["beta"] for lv in out["levels"]]
    assert betas[-1] == 1.0 and np.all(np.diff(betas) > 0)
    assert abs(out["log_evidence"] + math.log(400.0)) < 0.15
    post = np.stack([out["posteriors"]["u"].numpy(), out["posteriors"]["v"].numpy()], 1)
    np.testing.assert_allclose(post.mean(0), mean, atol=0.03)
    np.testing.assert_allclose(np.cov(post.T), cov, rtol=0.15, atol=3e-3)
    assert all(lv["acceptance"] > 0.3 for lv in out["levels"])


def test_jacobi_eigh_matches_the_solver_library():
    rng = np.random.default_rng(9)
    B, d = 6, 11
    A = rng.normal(size=(B, d, d))
    A = A + A.transpose(0, 2, 1)
    A[0] = np.diag(np.arange(1.0, d + 1.0))                     # already diagonal
    Q, _ = np.linalg.qr(rng.normal(size=(d, d)))
    A[1] = (Q * np.array([5.0, 5.0, 5.0] + list(range(8)))) @ Q.T  # repeated eigenvalues
    A[2] = (Q * np.logspace(-8, 8, d)) @ Q.T                    # wide dynamic range
    At = torch.as_tensor(A)
    lam, V = nuts.jacobi_eigh(At)
    ref = np.linalg.eigvalsh(A)
    for b in range(B):
        np.testing.assert_allclose(np.sort(lam[b].numpy()), ref[b], rtol=1e-12, atol=1e-12 * np.abs(ref[b]).max())
    np.testing.assert_allclose(((V * lam[:, None, :]) @ V.transpose(1, 2)).numpy(), A, rtol=0,
                               atol=1e-12 * np.abs(A).max())
    np.testing.assert_allclose((V.transpose(1, 2) @ V).numpy(), np.broadcast_to(np.eye(d), (B, d, d)), atol=1e-13)


def test_softabs_factors_capturable_and_library_agree():
    rng = np.random.default_rng(10)
    d = 7
    H = rng.normal(size=(4, d, d)) * 100.0
    H = torch.as_tensor(H + H.transpose(0, 2, 1))
    a1, f1 = nuts.softabs_factors(H, 0.3)
    a2, f2 = nuts.softabs_factors(H, 0.3, capturable=True)
    np.testing.assert_allclose(a2.numpy(), a1.numpy(), rtol=0, atol=1e-12 * float(a1.abs().max()))
    M = torch.linalg.inv(a1)
    for f in (f1, f2):  # F F^T = SoftAbs = inv(A)
        np.testing.assert_allclose((f @ f.transpose(1, 2)).numpy(), M.numpy(), rtol=0, atol=1e-10 * float(M.abs().max()))


def test_fixed_length_transition_equals_the_early_exit_one():
    """On the same random input the fixed-length (capturable) transition, which runs every
    doubling and leaf with finished trajectories masked, gives bit for bit the early-exiting
    transition's result."""
    rng = np.random.default_rng(11)
    d, P = 4, 64
    Q, _ = np.linalg.qr(rng.normal(size=(d, d)))
    cov = (Q * np.array([1.0, 0.3, 0.05, 2.0])) @ Q.T
    f = _gauss(np.zeros(d), cov)
    gen = torch.Generator().manual_seed(4)
    z = torch.as_tensor(rng.normal(size=(P, d)))
    logp, grad = f(z)
    inv_mass = torch.as_tensor(np.broadcast_to(cov, (P, d, d)).copy())
    _, F = nuts.softabs_factors(-torch.linalg.inv(inv_mass), 1e6)  # ~ |H|: F F^T = inv(cov)
    inv_mass_s = torch.linalg.inv(F @ F.transpose(1, 2))
    rand = nuts.nuts_draw(P, d, 6, gen, "cpu")
    a = nuts.nuts_step(f, z, logp, grad, inv_mass_s, 0.4, gen, 6, mass_factor=F, rand=rand)
    b = nuts.nuts_step(f, z, logp, grad, inv_mass_s, 0.4, gen, 6, mass_factor=F, rand=rand, fixed_length=True)
    for x, y in zip(a[:3], b[:3]):
        assert torch.equal(x, y)
    for k in a[3]:
        assert torch.equal(a[3][k], b[3][k]), k
    assert int(a[3]["num_integration_steps"].max()) < 2 ** 6 - 1  # some trajectories ended early
