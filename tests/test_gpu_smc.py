"""Tempered SMC with the HIP likelihood: levels progress, and one level (reweight, evidence,
resample, refit, IRMH rounds with the fused kernel) captured in a HIP graph replays exactly
like the eager level on the same random inputs, replay after replay."""

import math

import numpy as np
import pytest

from tests._setup import INJ, add_to_core, make_detector

pytestmark = pytest.mark.gpu

MODES = ((2, 2),)
# the box prior below is ~1e-4 wide in the angles and the mass ratio (variances ~1e-8): the
# reference's default regularization 1e-6 would put every proposal outside it
REG = 1e-14


def _likelihood():
    from pespace_amd.core import LikelihoodCore
    d, labels = make_detector(modes=MODES, T=131072.0 * 5, dt=5.0, seed=11)
    core = LikelihoodCore(modes=["22"], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    add_to_core(core, d, labels)

    class _Lk:
        pass

    lk = _Lk()
    lk.core = core
    lk.precision = "fp64"
    lk.log_likelihood_and_grad = lambda rows: core.lnl(rows, grad=True)
    return lk


def _prior(width=2e-4):
    """A narrow box around the injection for every parameter (the posterior of this short,
    loud signal is narrower still, so tempering takes several levels)."""
    pri = {}
    for k, v in INJ.items():
        h = width * max(abs(v), 1.0)
        pri[k] = {"distribution": "Uniform", "low": v - h, "high": v + h}
    return pri


def _clone(state):
    return {k: v.clone() for k, v in state.items()}


def test_smc_levels_with_hip_likelihood():
    import torch

    from pespace_amd.inference.smc import TemperedSMC, log_ess
    lk = _likelihood()
    s = TemperedSMC(lk, _prior(), num_particles=512, target_ess=0.9, num_mcmc_steps=4, rkey_seed=1,
                    covariance_regularization=REG)
    state = s.init_state()
    assert torch.all(torch.isfinite(state["logl"]))
    betas = []
    for _ in range(3):
        logl0, beta0 = state["logl"].clone(), state["beta"].clone()
        s.level(state, s.draw())
        d = float(state["delta"])
        betas.append(float(state["beta"]))
        assert d > 0.0 and float(state["beta"]) == pytest.approx(float(beta0) + d)
        # the increment puts the ESS of the incremental weights at the target (or takes the rest)
        le = float(log_ess(torch.nan_to_num(d * logl0)))
        assert betas[-1] == 1.0 or abs(le - math.log(512 * 0.9)) <= 1e-4
        assert 0.0 < float(state["accept"]) <= 1.0
        assert torch.all(torch.isfinite(state["logl"]))
    assert np.all(np.diff(betas) > 0)


def test_smc_level_hip_graph_replays_match_eager():
    import torch

    from pespace_amd.inference.smc import TemperedSMC
    lk = _likelihood()
    s = TemperedSMC(lk, _prior(), num_particles=256, target_ess=0.9, num_mcmc_steps=3, rkey_seed=2,
                    covariance_regularization=REG)
    state = s.init_state()
    s.level(state, s.draw())  # one eager level: sizes every buffer, moves beta off 0
    lk.core.pin_scratch(True)  # the captured level keeps the context's scratch pointers
    static_state = _clone(state)
    rand = s.draw()
    static_rand = {k: v.clone() for k, v in rand.items()}
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up on the capture stream
        s.level(_clone(state), static_rand)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        s.level(static_state, static_rand)
    for rep in range(3):
        start = _clone(state)
        r = s.draw()
        for k in static_state:
            static_state[k].copy_(start[k])
        for k in static_rand:
            static_rand[k].copy_(r[k])
        graph.replay()
        torch.cuda.synchronize()
        eager = _clone(start)
        s.level(eager, r)
        torch.cuda.synchronize()
        for k in ("x", "logl", "logprior", "beta", "log_evidence", "accept", "delta"):
            assert torch.equal(static_state[k], eager[k]), (rep, k)
        state = eager
    lk.core.pin_scratch(False)


def test_smc_with_nuts_mutation_on_hip_likelihood():
    """Two levels of the NUTS-mutation driver: per-particle SoftAbs metrics from the adjoint
    gradient's differences are symmetric positive definite, the trajectories take finite
    steps without divergences, and the tempering advances."""
    import torch

    from pespace_amd.inference.nuts import TemperedSMCNUTS
    lk = _likelihood()
    s = TemperedSMCNUTS(lk, _prior(), num_particles=256, target_ess=0.9, num_mcmc_steps=1, max_num_doublings=5,
                        rkey_seed=3)
    state = s.init_state()
    A = s.inverse_mass_matrices(state["x"][:32], torch.tensor(0.5, dtype=torch.float64, device="cuda"))
    assert torch.allclose(A, A.transpose(1, 2), rtol=0, atol=1e-12 * float(A.abs().max()))
    assert bool((torch.linalg.eigvalsh(A) > 0).all())
    betas = []
    for _ in range(2):
        s.level(state)
        betas.append(float(state["beta"]))
        info = s.last_info[-1]
        assert int(info["num_integration_steps"].min()) >= 1
        assert float(info["is_divergent"].float().mean()) < 0.05
        assert 0.2 < float(state["accept"]) <= 1.0
        assert torch.all(torch.isfinite(state["logl"]))
    assert betas[1] > betas[0] > 0.0


def test_nuts_level_hip_graph_replays_match_eager():
    """A capturable NUTS level (Jacobi metrics, fixed-length trajectories; the metric batch of
    2 d rows per particle and every leapfrog step on the fused lnL + gradient kernel, the
    z-gradients through torch autograd) captured once and replayed: equal to the eager level
    on the same random input, replay after replay."""
    import torch

    from pespace_amd.inference.nuts import TemperedSMCNUTS
    lk = _likelihood()
    s = TemperedSMCNUTS(lk, _prior(), num_particles=64, target_ess=0.9, num_mcmc_steps=1, max_num_doublings=3,
                        rkey_seed=4, capturable=True)
    state = s.init_state()
    print("nuts capture: eager level", flush=True)
    s.level(state, s.draw())  # sizes every buffer (the metric batch is the largest), moves beta off 0
    lk.core.pin_scratch(True)
    static_state = _clone(state)
    static_rand = {k: v.clone() for k, v in s.draw().items()}
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up on the capture stream
        s.level(_clone(state), static_rand)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    print("nuts capture: capturing", flush=True)
    with torch.cuda.graph(graph):
        s.level(static_state, static_rand)
    for rep in range(3):
        start = _clone(state)
        r = s.draw()
        for k in static_state:
            static_state[k].copy_(start[k])
        for k in static_rand:
            static_rand[k].copy_(r[k])
        graph.replay()
        torch.cuda.synchronize()
        eager = _clone(start)
        s.level(eager, r)
        torch.cuda.synchronize()
        for k in ("x", "logl", "logprior", "beta", "log_evidence", "accept", "delta"):
            assert torch.equal(static_state[k], eager[k]), (rep, k)
        assert torch.all(torch.isfinite(eager["logl"]))
        state = eager
    lk.core.pin_scratch(False)
