"""Adaptive-tempered SMC with a NUTS mutation on the unconstrained space, per-particle
SoftAbs-Hessian metrics, every density and gradient from the fused lnL + gradient kernel.

Follows pespace/inference/sampler/blackjax_smc.py ``BlackjaxSMCwithNUTS`` (:556-851):
  * sampling coordinates z: each sampled parameter with a bounded support is mapped by
    x = low + (high - low) sigmoid(z) (numpyro's biject_to for an interval), unbounded ones
    by the identity; the log-Jacobian is part of the prior term and is not tempered
    (sampler/base.py:507-523);
  * at every level, after resampling, each particle's dense inverse mass matrix is
    inv(SoftAbs(-H)) with H the Hessian of the tempered log-posterior at the particle
    (``_tune_mcmc_parameters`` :700-733; SoftAbs eigenvalues lambda / tanh(alpha lambda),
    ``_softabs_metric`` :116-128), frozen along the trajectory;
  * ``num_mcmc_steps`` NUTS transitions with a fixed shared step size, at most
    ``max_num_doublings`` doublings and a divergence threshold on the energy error.
The NUTS transition restates blackjax's (absent here): velocity-Verlet leapfrog with kinetic
energy p^T A p / 2 (A = inverse mass matrix, momentum p ~ N(0, A^-1)), trajectory doubling in
a random direction, multinomial sampling (uniform-progressive inside a subtree,
biased-progressive when merging it), the generalized U-turn criterion v_left . rho <= 0 or
v_right . rho <= 0 with rho = sum p - (p_left + p_right) / 2, checked inside the new subtree
at the checkpoints of the iterative scheme (leaf n: checkpoints popcount(n >> 1) - k for the
trailing one-bits of n) and over the whole trajectory after each merge.

All particles advance in lockstep: the subtree of doubling d has 2^d leaves for every
particle still building, so one leapfrog step is one batched density-and-gradient evaluation
of the whole cloud (finished particles are carried along and masked), as jax's vmap of the
per-particle kernel does.  The Hessians come from central differences of the exact adjoint
gradient in z (2 dim evaluations per particle in one batch).
"""

from __future__ import annotations

import math

import torch

from .smc import TemperedSMC, systematic_resample, tempering_increment


# ------------------------------------------------------------------------------------------
# unconstrained coordinates
# ------------------------------------------------------------------------------------------

def _support(dist):
    k, kw = dist.kind, dist.kw
    if k in ("Uniform", "LogUniform"):
        return kw["low"], kw["high"]
    if k == "Sine":
        return 0.0, math.pi
    if k == "Cosine":
        return -0.5 * math.pi, 0.5 * math.pi
    return None  # Normal: the real line


class Unconstrained:
    """x(z) per sampled parameter and log |dx/dz|, for a Prior."""

    def __init__(self, prior):
        self.prior = prior
        self.bounds = [_support(d) for d in prior.sampled.values()]

    def to_x(self, z):
        cols = []
        for i, b in enumerate(self.bounds):
            cols.append(z[:, i] if b is None else b[0] + (b[1] - b[0]) * torch.sigmoid(z[:, i]))
        return torch.stack(cols, 1)

    def to_z(self, x):
        cols = []
        for i, b in enumerate(self.bounds):
            if b is None:
                cols.append(x[:, i])
            else:
                u = (x[:, i] - b[0]) / (b[1] - b[0])
                cols.append(torch.log(u) - torch.log1p(-u))
        return torch.stack(cols, 1)

    def log_det(self, z):
        out = torch.zeros(z.shape[0], dtype=z.dtype, device=z.device)
        for i, b in enumerate(self.bounds):
            if b is not None:
                zi = z[:, i]
                out = out + math.log(b[1] - b[0]) - torch.nn.functional.softplus(-zi) \
                    - torch.nn.functional.softplus(zi)
        return out


# ------------------------------------------------------------------------------------------
# SoftAbs metric
# ------------------------------------------------------------------------------------------

def softabs_inverse_metric(neg_hessian, alpha):
    """inv(V diag(l / tanh(alpha l)) V^T) = V diag(tanh(alpha l) / l) V^T for a batch of
    symmetric matrices (l -> 0: the finite limit alpha)."""
    h = 0.5 * (neg_hessian + neg_hessian.transpose(-1, -2))
    lam, V = torch.linalg.eigh(h)
    al = alpha * lam
    small = torch.abs(al) < 1e-8
    inv_soft = torch.where(small, torch.full_like(lam, alpha), torch.tanh(al) / torch.where(small, 1.0, lam))
    return (V * inv_soft[..., None, :]) @ V.transpose(-1, -2)


def _bmv(A, v):
    return torch.einsum("pij,pj->pi", A, v)


def _is_turning(A, p_left, p_right, p_sum):
    rho = p_sum - 0.5 * (p_left + p_right)
    return ((_bmv(A, p_left) * rho).sum(1) <= 0.0) | ((_bmv(A, p_right) * rho).sum(1) <= 0.0)


def _logaddexp(a, b):
    return torch.logaddexp(a, b)


# ------------------------------------------------------------------------------------------
# NUTS transition (batched, lockstep)
# ------------------------------------------------------------------------------------------

def nuts_step(logp_grad, z, logp, grad, inv_mass, step_size, gen, max_num_doublings=7, divergence_threshold=100.0):
    """One NUTS transition for every particle.

    logp_grad(z) -> (log density [P], gradient [P, d]); inv_mass [P, d, d] (the kinetic
    energy's matrix A).  Returns (z, logp, grad, info) with info = dict(acceptance_rate,
    num_integration_steps, is_divergent) per particle.
    """
    P, d = z.shape
    dev, dt = z.device, z.dtype
    mass = torch.linalg.inv(inv_mass)
    mass = 0.5 * (mass + mass.transpose(-1, -2))
    Lm = torch.linalg.cholesky(mass)
    p0 = _bmv(Lm, torch.randn((P, d), generator=gen, device=dev, dtype=dt))
    kin = lambda p: 0.5 * (p * _bmv(inv_mass, p)).sum(1)  # noqa: E731
    H0 = -logp + kin(p0)

    # trajectory state
    ql, pl, gl, lpl = z.clone(), p0.clone(), grad.clone(), logp.clone()
    qr, pr, gr, lpr = z.clone(), p0.clone(), grad.clone(), logp.clone()
    qprop, lpprop, gprop = z.clone(), logp.clone(), grad.clone()
    log_w = torch.zeros(P, dtype=dt, device=dev)          # log sum of weights, weight = exp(H0 - H)
    p_sum = p0.clone()
    done = ~torch.isfinite(logp)
    diverged = torch.zeros(P, dtype=torch.bool, device=dev)
    n_steps = torch.zeros(P, dtype=torch.long, device=dev)
    acc_sum = torch.zeros(P, dtype=dt, device=dev)
    ck_p = torch.zeros((P, max_num_doublings + 1, d), dtype=dt, device=dev)
    ck_s = torch.zeros_like(ck_p)

    for depth in range(max_num_doublings):
        if bool(done.all()):
            break
        go_right = torch.rand(P, generator=gen, device=dev) < 0.5
        eps = torch.where(go_right, step_size, -step_size).to(dt)[:, None]
        q = torch.where(go_right[:, None], qr, ql)
        p = torch.where(go_right[:, None], pr, pl)
        g = torch.where(go_right[:, None], gr, gl)
        building = ~done
        s_log_w = torch.full((P,), -math.inf, dtype=dt, device=dev)
        s_sum = torch.zeros((P, d), dtype=dt, device=dev)
        s_q, s_lp, s_g = q.clone(), torch.full((P,), -math.inf, dtype=dt, device=dev), g.clone()
        s_turn = torch.zeros(P, dtype=torch.bool, device=dev)
        s_div = torch.zeros(P, dtype=torch.bool, device=dev)
        last_q, last_p, last_g, last_lp = q, p, g, torch.zeros(P, dtype=dt, device=dev)
        for n in range(2 ** depth):
            act = building & ~s_turn & ~s_div
            if not bool(act.any()):
                break
            ph = p + 0.5 * eps * g
            qn = q + eps * _bmv(inv_mass, ph)
            lpn, gn = logp_grad(qn)
            gn = torch.where(torch.isfinite(lpn)[:, None], gn, torch.zeros_like(gn))
            pn = ph + 0.5 * eps * gn
            Hn = -lpn + kin(pn)
            dH = Hn - H0
            dH = torch.where(torch.isnan(dH), torch.full_like(dH, math.inf), dH)
            div = act & (dH > divergence_threshold)
            lw_leaf = -dH
            n_steps = n_steps + act.long()
            acc_sum = acc_sum + torch.where(act, torch.clamp(torch.exp(-dH), max=1.0), torch.zeros_like(dH))
            # uniform progressive sampling inside the subtree
            new_s_log_w = torch.where(act, _logaddexp(s_log_w, lw_leaf), s_log_w)
            u = torch.rand(P, generator=gen, device=dev, dtype=dt)
            take = act & (torch.log(u) < lw_leaf - new_s_log_w)
            s_q = torch.where(take[:, None], qn, s_q)
            s_lp = torch.where(take, lpn, s_lp)
            s_g = torch.where(take[:, None], gn, s_g)
            s_log_w = new_s_log_w
            s_sum = torch.where(act[:, None], s_sum + pn, s_sum)
            # iterative U-turn checks inside the subtree
            if n % 2 == 0:
                idx = bin(n >> 1).count("1")
                ck_p[:, idx] = torch.where(act[:, None], pn, ck_p[:, idx])
                ck_s[:, idx] = torch.where(act[:, None], s_sum, ck_s[:, idx])
            else:
                idx_max = bin(n >> 1).count("1")
                n_sub = bin((~n & (n + 1)) - 1).count("1")
                turn = torch.zeros(P, dtype=torch.bool, device=dev)
                for i in range(idx_max, idx_max - n_sub, -1):
                    sub_sum = s_sum - ck_s[:, i] + ck_p[:, i]
                    turn = turn | _is_turning(inv_mass, ck_p[:, i], pn, sub_sum)
                s_turn = s_turn | (act & turn)
            s_div = s_div | div
            q = torch.where(act[:, None], qn, q)
            p = torch.where(act[:, None], pn, p)
            g = torch.where(act[:, None], gn, g)
            last_q, last_p, last_g = q, p, g
            last_lp = torch.where(act, lpn, last_lp)
        # merge the subtree where it neither turned nor diverged
        ok = building & ~s_turn & ~s_div
        diverged = diverged | (building & s_div)
        u = torch.rand(P, generator=gen, device=dev, dtype=dt)
        take = ok & (torch.log(u) < s_log_w - log_w)
        qprop = torch.where(take[:, None], s_q, qprop)
        lpprop = torch.where(take, s_lp, lpprop)
        gprop = torch.where(take[:, None], s_g, gprop)
        log_w = torch.where(ok, _logaddexp(log_w, s_log_w), log_w)
        p_sum = torch.where(ok[:, None], p_sum + s_sum, p_sum)
        right = ok & go_right
        left = ok & ~go_right
        qr = torch.where(right[:, None], last_q, qr)
        pr = torch.where(right[:, None], last_p, pr)
        gr = torch.where(right[:, None], last_g, gr)
        ql = torch.where(left[:, None], last_q, ql)
        pl = torch.where(left[:, None], last_p, pl)
        gl = torch.where(left[:, None], last_g, gl)
        turn_all = _is_turning(inv_mass, pl, pr, p_sum)
        done = done | (building & (s_turn | s_div)) | (ok & turn_all)
    info = {"acceptance_rate": acc_sum / torch.clamp(n_steps, min=1).to(dt), "num_integration_steps": n_steps,
            "is_divergent": diverged}
    return qprop, lpprop, gprop, info


# ------------------------------------------------------------------------------------------
# sampler
# ------------------------------------------------------------------------------------------

class TemperedSMCNUTS(TemperedSMC):
    """Adaptive-tempered SMC with NUTS mutations (BlackjaxSMCwithNUTS's algorithm).

    likelihood: an object with ``log_likelihood_and_grad(rows) -> (lnL [P], d lnL / d rows
    [P, n])`` (WhittleLikelihoodHip).  Particles live in the unconstrained coordinates z;
    ``run`` returns physical posteriors as TemperedSMC does.
    """

    def __init__(self, likelihood, prior_dict, num_particles=4096, target_ess=0.9, num_mcmc_steps=1,
                 step_size=0.2, softabs_alpha=0.3, max_num_doublings=7, divergence_threshold=100,
                 hessian_step=1e-4, rkey_seed=0, device=None, names=None):
        super().__init__(likelihood, prior_dict, num_particles=num_particles, target_ess=target_ess,
                         num_mcmc_steps=num_mcmc_steps, rkey_seed=rkey_seed, device=device, names=names)
        if step_size <= 0.0:
            raise ValueError("step_size must be positive.")
        self.step_size = float(step_size)
        self.softabs_alpha = float(softabs_alpha)
        self.max_num_doublings = int(max_num_doublings)
        self.divergence_threshold = float(divergence_threshold)
        self.hessian_step = float(hessian_step)
        self.bij = Unconstrained(self.prior)
        self.n_grad_evals = 0

    # log prior (with the Jacobian) and log likelihood with their z-gradients
    def _prior_and_like(self, z):
        with torch.enable_grad():
            zz = z.detach().requires_grad_(True)
            x = self.bij.to_x(zz)
            lp = self.prior.log_prob(x) + self.bij.log_det(zz)
            rows = self.prior.rows(x)
            lnl, g_rows = self.likelihood.log_likelihood_and_grad(rows.detach())
            lnl = lnl.to(z.dtype)
            g_rows = torch.where(torch.isfinite(lnl)[:, None], g_rows.to(z.dtype), torch.zeros_like(g_rows))
            g_lp, = torch.autograd.grad(lp.sum(), zz, retain_graph=True)
            g_ll, = torch.autograd.grad((rows * g_rows).sum(), zz)
        self.n_grad_evals += z.shape[0]
        return lp.detach(), lnl.detach(), g_lp, g_ll

    def _logpost_grad(self, beta):
        def f(z):
            lp, ll, g_lp, g_ll = self._prior_and_like(z)
            ok = torch.isfinite(ll)
            val = torch.where(ok, lp + beta * ll, torch.full_like(lp, -math.inf))
            return val, g_lp + beta * g_ll
        return f

    def init_state(self):
        x = self.prior.sample(self.num_particles, self.gen, self.device)
        z = self.bij.to_z(x)
        lp, ll, _, _ = self._prior_and_like(z)
        self.n_like_evals += self.num_particles
        zero = torch.zeros((), dtype=torch.float64, device=self.device)
        return {"x": z, "logl": ll, "logprior": lp, "beta": zero.clone(), "log_evidence": zero.clone(),
                "accept": zero.clone(), "delta": zero.clone()}

    def inverse_mass_matrices(self, z, beta):
        """inv(SoftAbs(-H)) of the tempered log-posterior at every particle (central
        differences of the exact gradient, all 2 d shifted points in one batch)."""
        P, d = z.shape
        h = self.hessian_step
        shifts = torch.eye(d, dtype=z.dtype, device=z.device) * h
        zz = torch.cat([(z[:, None, :] + shifts[None]).reshape(-1, d), (z[:, None, :] - shifts[None]).reshape(-1, d)])
        _, g = self._logpost_grad(beta)(zz)
        gp, gm = g[:P * d].reshape(P, d, d), g[P * d:].reshape(P, d, d)
        H = (gp - gm) / (2.0 * h)  # H[p, k, :] = d grad / d z_k
        H = torch.nan_to_num(H, nan=0.0, posinf=0.0, neginf=0.0)
        return softabs_inverse_metric(-H, self.softabs_alpha)

    def level(self, state, rand=None):
        z, logl, logprior, beta = state["x"], state["logl"], state["logprior"], state["beta"]
        u = (rand or self.draw())["u"]
        delta = tempering_increment(logl, beta, self.target_ess)
        beta_new = beta + delta
        lw = torch.nan_to_num(delta * logl)
        state["log_evidence"].add_(torch.logsumexp(lw, 0) - math.log(self.num_particles))
        idx = systematic_resample(lw, u)
        z = z[idx]
        inv_mass = self.inverse_mass_matrices(z, beta_new)
        f = self._logpost_grad(beta_new)
        logp, grad = f(z)
        acc = torch.zeros((), dtype=torch.float64, device=self.device)
        self.last_info = []
        for _ in range(self.num_mcmc_steps):
            z, logp, grad, info = nuts_step(f, z, logp, grad, inv_mass, self.step_size, self.gen,
                                            self.max_num_doublings, self.divergence_threshold)
            acc = acc + info["acceptance_rate"].mean()
            self.last_info.append(info)
            self.n_like_evals += int(info["num_integration_steps"].sum())
        lp, ll, _, _ = self._prior_and_like(z)
        state["x"].copy_(z)
        state["logl"].copy_(ll)
        state["logprior"].copy_(lp)
        state["beta"].copy_(beta_new)
        state["accept"].copy_(acc / self.num_mcmc_steps)
        state["delta"].copy_(delta)
        return state

    def draw(self):
        """Random input of one level that the caller may fix: the resampling offset (the
        momenta and the trajectory choices come from the sampler's generator)."""
        return {"u": torch.rand((), generator=self.gen, device=self.device, dtype=torch.float64)}

    def _physical(self, z):
        return self.bij.to_x(z)
