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

A level's random input (resampling offset, momenta, directions, leaf and merge uniforms) is
drawn up front (``draw``).  With ``capturable=True`` the level is a fixed sequence of device
operations: the metrics' eigen-decompositions by Jacobi rotations instead of the solver
library, and trajectories of fixed length (every doubling and leaf executed, finished
particles masked, no host synchronisation) -- the same transitions as the early-exiting
default on the same random input, at 2^max_num_doublings - 1 evaluations per particle -- so
that the whole level, likelihood launches included, can be captured in a HIP graph.
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

def jacobi_eigh(h, sweeps=12):
    """Eigenvalues and eigenvectors of a batch of small symmetric matrices [B, d, d] by cyclic
    Jacobi rotations (Press et al., Numerical Recipes, sec. 11.1: t = sgn(theta) / (|theta| +
    sqrt(theta^2 + 1)), theta = (a_qq - a_pp) / (2 a_pq)), a fixed number of sweeps over all
    pairs (p, q) written in elementwise device operations: no solver library call and no host
    synchronisation, so it can sit inside a captured graph.  Twelve sweeps take d = 11 to
    working precision (Jacobi converges quadratically once the off-diagonal part is small).
    Returns (lam [B, d], V [B, d, d]) with h = V diag(lam) V^T (unsorted)."""
    A = 0.5 * (h + h.transpose(-1, -2))
    d = A.shape[-1]
    V = torch.eye(d, dtype=A.dtype, device=A.device).expand(A.shape).clone()
    for _ in range(sweeps):
        for p in range(d - 1):
            for q in range(p + 1, d):
                apq = A[:, p, q]
                zero = apq == 0.0
                theta = (A[:, q, q] - A[:, p, p]) / (2.0 * torch.where(zero, torch.ones_like(apq), apq))
                t = torch.sign(theta) / (torch.abs(theta) + torch.sqrt(theta * theta + 1.0))
                t = torch.where(zero | ~torch.isfinite(t), torch.zeros_like(t), t)
                t = torch.where((theta == 0.0) & ~zero, torch.ones_like(t), t)  # sgn(0) = +1
                c = 1.0 / torch.sqrt(t * t + 1.0)
                sn = t * c
                c_, s_ = c[:, None], sn[:, None]
                colp, colq = A[:, :, p].clone(), A[:, :, q]
                A[:, :, p] = c_ * colp - s_ * colq
                A[:, :, q] = s_ * colp + c_ * colq
                rowp, rowq = A[:, p, :].clone(), A[:, q, :]
                A[:, p, :] = c_ * rowp - s_ * rowq
                A[:, q, :] = s_ * rowp + c_ * rowq
                vp, vq = V[:, :, p].clone(), V[:, :, q]
                V[:, :, p] = c_ * vp - s_ * vq
                V[:, :, q] = s_ * vp + c_ * vq
    return torch.diagonal(A, dim1=-2, dim2=-1).clone(), V


def softabs_factors(neg_hessian, alpha, capturable=False):
    """SoftAbs metric of a batch of symmetric matrices: with eigenvalues l and vectors V,
    M = V diag(l / tanh(alpha l)) V^T (l -> 0: the finite limit 1 / alpha).  Returns
    (inv(M) = V diag(tanh(alpha l) / l) V^T, F = V diag(sqrt(l / tanh(alpha l)))), F F^T = M:
    the kinetic energy's matrix and a factor that draws momenta from N(0, M).  capturable:
    the eigen-decomposition by Jacobi rotations (jacobi_eigh) instead of the solver library."""
    h = 0.5 * (neg_hessian + neg_hessian.transpose(-1, -2))
    lam, V = jacobi_eigh(h) if capturable else torch.linalg.eigh(h)
    al = alpha * lam
    small = torch.abs(al) < 1e-8
    inv_soft = torch.where(small, torch.full_like(lam, alpha), torch.tanh(al) / torch.where(small, 1.0, lam))
    inv_m = (V * inv_soft[..., None, :]) @ V.transpose(-1, -2)
    return inv_m, V * torch.rsqrt(inv_soft)[..., None, :]


def softabs_inverse_metric(neg_hessian, alpha):
    """inv(V diag(l / tanh(alpha l)) V^T) = V diag(tanh(alpha l) / l) V^T for a batch of
    symmetric matrices (l -> 0: the finite limit alpha)."""
    return softabs_factors(neg_hessian, alpha)[0]


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

def nuts_draw(P, d, max_num_doublings, gen, device, dtype=torch.float64):
    """The random input of one transition, drawn up front: momentum normals [P, d], the
    direction of each doubling [max, P], the uniform of every leaf [2^max - 1, P] (leaf n of
    doubling depth is row 2^depth - 1 + n) and of every subtree merge [max, P]."""
    return {"p": torch.randn((P, d), generator=gen, device=device, dtype=dtype),
            "dir": torch.rand((max_num_doublings, P), generator=gen, device=device, dtype=dtype),
            "leaf": torch.rand((2 ** max_num_doublings - 1, P), generator=gen, device=device, dtype=dtype),
            "merge": torch.rand((max_num_doublings, P), generator=gen, device=device, dtype=dtype)}


def nuts_step(logp_grad, z, logp, grad, inv_mass, step_size, gen, max_num_doublings=7, divergence_threshold=100.0,
              mass_factor=None, rand=None, fixed_length=False):
    """One NUTS transition for every particle.

    logp_grad(z) -> (log density [P], gradient [P, d]); inv_mass [P, d, d] (the kinetic
    energy's matrix A); mass_factor [P, d, d]: F with F F^T = inv(A) (by default from a
    Cholesky factorisation of the inverse); rand: the transition's random input (nuts_draw;
    by default drawn from gen as the trajectory grows).  fixed_length: every doubling and every
    leaf is executed, finished trajectories under masks, with no host synchronisation (the
    same transition as the default's early exits, for a captured graph: 2^max - 1 likelihood
    evaluations per particle).  Returns (z, logp, grad, info) with info =
    dict(acceptance_rate, num_integration_steps, is_divergent) per particle.
    """
    P, d = z.shape
    dev, dt = z.device, z.dtype
    if mass_factor is None:
        mass = torch.linalg.inv(inv_mass)
        mass = 0.5 * (mass + mass.transpose(-1, -2))
        mass_factor = torch.linalg.cholesky(mass)
    normals = rand["p"] if rand is not None else torch.randn((P, d), generator=gen, device=dev, dtype=dt)
    p0 = _bmv(mass_factor, normals)
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
        if not fixed_length and bool(done.all()):
            break
        go_right = (rand["dir"][depth] if rand is not None else torch.rand(P, generator=gen, device=dev)) < 0.5
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
            if not fixed_length and not bool(act.any()):
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
            u = rand["leaf"][2 ** depth - 1 + n] if rand is not None else torch.rand(P, generator=gen, device=dev,
                                                                                   dtype=dt)
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
        u = rand["merge"][depth] if rand is not None else torch.rand(P, generator=gen, device=dev, dtype=dt)
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
                 hessian_step=1e-4, rkey_seed=0, device=None, names=None, capturable=False):
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
        # capturable: a level is a fixed sequence of device operations (Jacobi eigen-
        # decompositions for the metrics, fixed-length trajectories), so that it can be captured
        # once in a HIP graph and replayed; each transition then costs 2^max_num_doublings - 1
        # likelihood and gradient evaluations per particle
        self.capturable = bool(capturable)

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
        return self._metrics(z, beta)[0]

    def _metrics(self, z, beta):
        """(inv(SoftAbs(-H)), a factor F of SoftAbs(-H) = F F^T) at every particle."""
        P, d = z.shape
        h = self.hessian_step
        shifts = torch.eye(d, dtype=z.dtype, device=z.device) * h
        zz = torch.cat([(z[:, None, :] + shifts[None]).reshape(-1, d), (z[:, None, :] - shifts[None]).reshape(-1, d)])
        _, g = self._logpost_grad(beta)(zz)
        gp, gm = g[:P * d].reshape(P, d, d), g[P * d:].reshape(P, d, d)
        H = (gp - gm) / (2.0 * h)  # H[p, k, :] = d grad / d z_k
        H = torch.nan_to_num(H, nan=0.0, posinf=0.0, neginf=0.0)
        return softabs_factors(-H, self.softabs_alpha, capturable=self.capturable)

    def level(self, state, rand=None):
        z, logl, logprior, beta = state["x"], state["logl"], state["logprior"], state["beta"]
        rand = rand if rand is not None else self.draw()
        u = rand["u"]
        delta = tempering_increment(logl, beta, self.target_ess)
        beta_new = beta + delta
        lw = torch.nan_to_num(delta * logl)
        state["log_evidence"].add_(torch.logsumexp(lw, 0) - math.log(self.num_particles))
        idx = systematic_resample(lw, u)
        z = z[idx]
        inv_mass, mass_factor = self._metrics(z, beta_new)
        f = self._logpost_grad(beta_new)
        logp, grad = f(z)
        acc = torch.zeros((), dtype=torch.float64, device=self.device)
        self.last_info = []
        for k in range(self.num_mcmc_steps):
            tr = {key: rand[key][k] for key in ("p", "dir", "leaf", "merge")}
            z, logp, grad, info = nuts_step(f, z, logp, grad, inv_mass, self.step_size, self.gen,
                                            self.max_num_doublings, self.divergence_threshold,
                                            mass_factor=mass_factor, rand=tr, fixed_length=self.capturable)
            acc = acc + info["acceptance_rate"].mean()
            self.last_info.append(info)
            if self.capturable:  # every leaf of every doubling is evaluated
                self.n_like_evals += self.num_particles * (2 ** self.max_num_doublings - 1)
            else:
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
        """Random input of one level, drawn up front: the resampling offset and, per
        transition (leading axis num_mcmc_steps), the momenta's normals and the trajectory's
        uniforms (nuts_draw)."""
        u = torch.rand((), generator=self.gen, device=self.device, dtype=torch.float64)
        tr = [nuts_draw(self.num_particles, self.prior.dim, self.max_num_doublings, self.gen, self.device)
              for _ in range(self.num_mcmc_steps)]
        return {"u": u, **{key: torch.stack([t[key] for t in tr]) for key in ("p", "dir", "leaf", "merge")}}

    def _physical(self, z):
        return self.bij.to_x(z)
