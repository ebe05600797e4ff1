"""Adaptive-tempered SMC around the batched HIP likelihood (independent Metropolis-Hastings
mutation), with one tempering level capturable in a HIP graph.

Follows pespace/inference/sampler/blackjax_smc.py:
  * one level (``_TunedAdaptiveTemperedSMC._temper``, :250-307): the tempering increment
    that brings the effective sample size of the incremental weights to
    ``target_ess * num_particles``, found by bisection on [0, 1 - beta] over the current
    log-likelihoods; the log-evidence increment logsumexp(w) - log N; systematic
    resampling; the proposal refit;
  * the mutation (``BlackjaxSMCwithIRMH``, :854-1032): ``num_mcmc_steps`` rounds of
    independent Metropolis-Hastings in the physical space with the Gaussian proposal
    N(mean, cov + covariance_regularization I) fitted to the resampled particles
    (``_fitted_proposal``, :1025-1032), targeting log prior + beta log L;
  * the driver (``_inference_loop``, :334-436): levels until beta reaches 1.
The blackjax primitives it calls (absent here) are restated from their published
definitions: log ESS = 2 logsumexp(w) - logsumexp(2 w); systematic resampling
idx_i = searchsorted(cumsum(W), (i + u) / N) clipped to N - 1; the particle covariance with
the 1 / (N - 1) normalisation; the bisection solver (stop when |f(mid)| <= 1e-4 or after 100
halvings, take max_delta when f(max_delta) >= 0).

Priors follow the reference's ``prior_dict`` (sampler/base.py:36-160, 453-600):
``{"distribution": "Uniform", "low": a, "high": b}`` (also LogUniform, Normal, Sine,
Cosine, Delta), ``{"fixed": v}``, ``{"transform": fn, "from": (names...)}`` (fn acts on
torch tensors), validated in declaration order.

All state lives on the likelihood's device.  ``level(state, rand)`` takes its random inputs
(the resampling offset, the proposal normals and the acceptance uniforms) as tensors and is a
fixed sequence of device operations without host synchronisation, so it can be captured once
and replayed (the likelihood's scratch is sized by a first eager call).
"""

from __future__ import annotations

import math
import time

import torch

from ..core import PARAM_NAMES

_LOG_2PI = math.log(2.0 * math.pi)


# ------------------------------------------------------------------------------------------
# priors
# ------------------------------------------------------------------------------------------

class Sine:
    """p(x) = sin(x) / 2 on [0, pi] (sampler/base.py:38-56)."""


class Cosine:
    """p(x) = cos(x) / 2 on [-pi/2, pi/2] (sampler/base.py:59-77)."""


class _Dist:
    """One sampled parameter: inverse-CDF sampling from uniforms and log density."""

    def __init__(self, kind, **kw):
        self.kind = kind
        self.kw = {k: float(v) for k, v in kw.items()}
        if kind in ("Uniform", "LogUniform"):
            lo, hi = self.kw["low"], self.kw["high"]
            if not hi > lo or (kind == "LogUniform" and lo <= 0.0):
                raise ValueError(f"invalid bounds for {kind}: low={lo}, high={hi}")
        elif kind == "Normal":
            if not self.kw["scale"] > 0.0:
                raise ValueError("Normal scale must be positive")
        elif kind not in ("Sine", "Cosine"):
            raise ValueError(f"unsupported distribution '{kind}'")

    def sample_from_uniform(self, u):
        k, kw = self.kind, self.kw
        if k == "Uniform":
            return kw["low"] + (kw["high"] - kw["low"]) * u
        if k == "LogUniform":
            return torch.exp(math.log(kw["low"]) + (math.log(kw["high"]) - math.log(kw["low"])) * u)
        if k == "Sine":
            return torch.arccos(1.0 - 2.0 * u)
        if k == "Cosine":
            return torch.arcsin(2.0 * u - 1.0)
        # Normal: inverse CDF through erfinv
        return kw["loc"] + kw["scale"] * math.sqrt(2.0) * torch.erfinv(2.0 * u - 1.0)

    def log_prob(self, x):
        k, kw = self.kind, self.kw
        ninf = torch.full_like(x, -math.inf)
        if k == "Uniform":
            inside = (x >= kw["low"]) & (x <= kw["high"])
            return torch.where(inside, torch.full_like(x, -math.log(kw["high"] - kw["low"])), ninf)
        if k == "LogUniform":
            inside = (x >= kw["low"]) & (x <= kw["high"])
            norm = math.log(math.log(kw["high"]) - math.log(kw["low"]))
            return torch.where(inside, -torch.log(torch.where(inside, x, torch.ones_like(x))) - norm, ninf)
        if k == "Sine":
            inside = (x >= 0.0) & (x <= math.pi)
            return torch.where(inside, torch.log(0.5 * torch.sin(x)), ninf)
        if k == "Cosine":
            inside = (x >= -0.5 * math.pi) & (x <= 0.5 * math.pi)
            return torch.where(inside, torch.log(0.5 * torch.cos(x)), ninf)
        z = (x - kw["loc"]) / kw["scale"]
        return -0.5 * z * z - math.log(kw["scale"]) - 0.5 * _LOG_2PI


class Prior:
    """A reference-style ``prior_dict`` over the likelihood's parameters.

    ``names`` are the likelihood's parameter names in row order (``PARAM_NAMES`` for
    WhittleLikelihoodHip).  Sampled parameters become the columns of the particle matrix in
    declaration order; ``rows(x)`` assembles the likelihood rows with the fixed and derived
    ones.  Raises ValueError with the reference's messages for undeclared inputs, duplicate
    declarations and missing likelihood parameters (sampler/base.py:100-160).
    """

    def __init__(self, prior_dict, names=PARAM_NAMES):
        self.names = tuple(names)
        self.sampled, self.fixed, self.derived = {}, {}, []
        declared = set()
        for name, spec in prior_dict.items():
            if not isinstance(spec, dict):
                raise ValueError(f"Prior entry for '{name}' must be a dict, got {type(spec).__name__}.")
            if "transform" in spec:
                if "from" not in spec:
                    raise ValueError(f"Derived parameter entry for '{name}' must provide a 'from' key listing "
                                     "its input parameter names.")
                outputs = (name,) if isinstance(name, str) else tuple(name)
                inputs = tuple(spec["from"])
                missing = [k for k in inputs if k not in declared]
                if missing:
                    raise ValueError(f"Derived parameter(s) {sorted(outputs)} depend on parameter(s) {missing} "
                                     "that are not declared before them.")
                overlap = sorted(set(outputs) & declared)
                if overlap:
                    raise ValueError(f"Derived parameter(s) {overlap} are declared more than once.")
                declared.update(outputs)
                self.derived.append((outputs, spec["transform"], inputs))
                continue
            if name in declared:
                raise ValueError(f"Parameter '{name}' is declared more than once (fixed and sampled/derived "
                                 "declarations must not conflict).")
            declared.add(name)
            if "fixed" in spec:
                self.fixed[name] = float(spec["fixed"])
            elif "distribution" in spec:
                kw = {k: v for k, v in spec.items() if k != "distribution"}
                d = spec["distribution"]
                kind = d if isinstance(d, str) else d.__name__
                if kind == "Delta":
                    self.fixed[name] = float(kw.get("v", kw.get("value")))
                else:
                    self.sampled[name] = _Dist(kind, **kw)
            else:
                raise ValueError(f"Prior entry for '{name}' must contain a 'distribution', 'transform' or "
                                 "'fixed' key.")
        missing = set(self.names) - declared
        if missing:
            raise ValueError("The prior does not provide all parameters required by the likelihood. Missing: "
                             f"{sorted(missing)}. Provided by the prior: {sorted(declared)}.")
        self.sampled_names = tuple(self.sampled)
        self.dim = len(self.sampled_names)

    def sample(self, n, generator, device, dtype=torch.float64):
        u = torch.rand((n, self.dim), generator=generator, device=device, dtype=dtype)
        return torch.stack([d.sample_from_uniform(u[:, i]) for i, d in enumerate(self.sampled.values())], dim=1)

    def log_prob(self, x):
        lp = torch.zeros(x.shape[0], dtype=x.dtype, device=x.device)
        for i, d in enumerate(self.sampled.values()):
            lp = lp + d.log_prob(x[:, i])
        return lp

    def rows(self, x, out=None):
        """Likelihood rows [n, len(names)] from particles x [n, dim]."""
        vals = {name: x[:, i] for i, name in enumerate(self.sampled_names)}
        for name, v in self.fixed.items():
            vals[name] = torch.full((x.shape[0],), v, dtype=x.dtype, device=x.device)
        for outputs, fn, inputs in self.derived:
            res = fn(**{k: vals[k] for k in inputs})
            if len(outputs) == 1:
                vals[outputs[0]] = res
            else:
                for o, r in zip(outputs, res):
                    vals[o] = r
        cols = [vals[k] for k in self.names]
        if out is None:
            return torch.stack(cols, dim=1)
        for j, c in enumerate(cols):
            out[:, j] = c
        return out


# ------------------------------------------------------------------------------------------
# SMC primitives (device tensors, no host synchronisation)
# ------------------------------------------------------------------------------------------

def log_ess(log_weights):
    """log of the effective sample size of unnormalised log-weights: 2 lse(w) - lse(2 w)."""
    return 2.0 * torch.logsumexp(log_weights, 0) - torch.logsumexp(2.0 * log_weights, 0)


def tempering_increment(loglik, beta, target_ess, eps=1e-4, max_iter=100):
    """Increment delta in [0, 1 - beta] with log ESS(delta loglik) = log(N target_ess).

    f(delta) = log ESS(nan_to_num(delta loglik)) - log(N target_ess) decreases from
    log(N / N_target) >= 0; if f(1 - beta) >= 0 the whole remaining step is taken, otherwise
    the bracket [0, 1 - beta] is halved towards the root until |f(mid)| <= eps or max_iter
    halvings (every halving is executed, with the bracket frozen once converged, so the
    number of device operations is fixed).  Returns a 0-d tensor.
    """
    n = loglik.shape[0]
    target = math.log(n * target_ess)

    def f(d):
        return log_ess(torch.nan_to_num(d * loglik)) - target

    hi = 1.0 - beta
    lo = torch.zeros_like(hi)
    f_hi = f(hi)
    mid = 0.5 * (lo + hi)
    done = torch.zeros((), dtype=torch.bool, device=loglik.device)
    for _ in range(max_iter):
        mid = torch.where(done, mid, 0.5 * (lo + hi))
        f_mid = f(mid)
        conv = torch.abs(f_mid) <= eps
        # f decreasing: root right of mid when f(mid) > 0
        go_right = f_mid > 0
        lo = torch.where(done | ~go_right, lo, mid)
        hi = torch.where(done | go_right, hi, mid)
        done = done | conv
    return torch.where(f_hi >= 0, 1.0 - beta, mid)


def systematic_resample(log_weights, u):
    """Indices of systematic resampling: searchsorted(cumsum(W), (i + u) / N), clipped."""
    n = log_weights.shape[0]
    w = torch.exp(log_weights - torch.logsumexp(log_weights, 0))
    cs = torch.cumsum(w, 0)
    pos = (torch.arange(n, dtype=w.dtype, device=w.device) + u) / n
    idx = torch.searchsorted(cs, pos)
    return torch.clamp(idx, max=n - 1)


def _cholesky_small(a):
    """Lower Cholesky factor of a small SPD matrix by columns, in elementwise device ops
    (no solver library call, so it can sit inside a captured graph)."""
    d = a.shape[0]
    L = torch.zeros_like(a)
    for j in range(d):
        v = a[j:, j] - (L[j:, :j] * L[j, :j]).sum(1) if j > 0 else a[j:, j].clone()
        ljj = torch.sqrt(torch.clamp(v[0], min=torch.finfo(a.dtype).tiny))
        L[j, j] = ljj
        if j + 1 < d:
            L[j + 1:, j] = v[1:] / ljj
    return L


def _lower_solve_rows(L, r):
    """z with L z_i = r_i for every row r_i of r [n, d] (forward substitution)."""
    d = L.shape[0]
    cols = []
    for j in range(d):
        acc = r[:, j]
        for k in range(j):
            acc = acc - L[j, k] * cols[k]
        cols.append(acc / L[j, j])
    return torch.stack(cols, 1)


def fit_gaussian(x, reg):
    """Mean, covariance (1 / (N - 1)) + reg I, and its lower Cholesky factor."""
    mean = x.mean(0)
    xc = x - mean
    cov = (xc[:, :, None] * xc[:, None, :]).sum(0) / (x.shape[0] - 1)
    cov = cov + reg * torch.eye(x.shape[1], dtype=x.dtype, device=x.device)
    return mean, cov, _cholesky_small(cov)


def gaussian_logpdf_chol(x, mean, chol):
    """log N(x; mean, L L^T) per row."""
    z = _lower_solve_rows(chol, x - mean)
    return -0.5 * (z * z).sum(1) - torch.log(torch.diagonal(chol)).sum() - 0.5 * x.shape[1] * _LOG_2PI


def gaussian_from_normals(mean, chol, z):
    """mean + L z_i for every row z_i of z [n, d]."""
    return mean + (z[:, None, :] * chol[None, :, :]).sum(2)


def _tempered(beta, logl):
    """beta log L, with -inf (outside the likelihood's domain) kept at -inf for any beta."""
    return torch.where(torch.isfinite(logl), beta * logl, torch.full_like(logl, -math.inf))


# ------------------------------------------------------------------------------------------
# sampler
# ------------------------------------------------------------------------------------------

class TemperedSMC:
    """Adaptive-tempered SMC with IRMH mutations (BlackjaxSMCwithIRMH's algorithm).

    likelihood: an object with ``log_likelihood_batch(rows) -> lnL`` on device tensors
    (WhittleLikelihoodHip, or any batched callable wrapped in such an object); when it has a
    ``core`` (LikelihoodCore), the mutation writes lnL into a preallocated buffer, which keeps
    a level capturable.
    """

    def __init__(self, likelihood, prior_dict, num_particles=4096, target_ess=0.9, num_mcmc_steps=100,
                 covariance_regularization=1e-6, rkey_seed=0, device=None, names=None):
        if num_mcmc_steps < 1:
            raise ValueError("num_mcmc_steps must be positive.")
        if not 0.0 < target_ess <= 1.0:
            raise ValueError("target_ess must lie in (0, 1].")
        self.likelihood = likelihood
        if names is None:
            names = likelihood.likelihood_param_type._fields if hasattr(likelihood, "likelihood_param_type") \
                else PARAM_NAMES
        self.prior = Prior(prior_dict, names)
        self.num_particles = int(num_particles)
        self.target_ess = float(target_ess)
        self.num_mcmc_steps = int(num_mcmc_steps)
        self.reg = float(covariance_regularization)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda" if torch.cuda.is_available() else "cpu")
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(int(rkey_seed))
        self._rows = None
        self._lnl = None
        self.n_like_evals = 0

    # likelihood of particles (into preallocated buffers when the core allows it)
    def _loglik(self, x):
        n = x.shape[0]
        if self._rows is None or self._rows.shape[0] != n:
            self._rows = torch.empty((n, len(self.prior.names)), dtype=torch.float64, device=x.device)
            self._lnl = torch.empty(n, dtype=torch.float64, device=x.device)
        rows = self.prior.rows(x, out=self._rows)
        core = getattr(self.likelihood, "core", None)
        if core is not None:
            core.lnl(rows, out=self._lnl, precision=getattr(self.likelihood, "precision", "fp64"))
            return self._lnl.clone()
        return self.likelihood.log_likelihood_batch(rows).to(torch.float64)

    def init_state(self):
        x = self.prior.sample(self.num_particles, self.gen, self.device)
        logl = self._loglik(x)
        self.n_like_evals += self.num_particles
        return {"x": x, "logl": logl, "logprior": self.prior.log_prob(x),
                "beta": torch.zeros((), dtype=torch.float64, device=self.device),
                "log_evidence": torch.zeros((), dtype=torch.float64, device=self.device),
                "accept": torch.zeros((), dtype=torch.float64, device=self.device),
                "delta": torch.zeros((), dtype=torch.float64, device=self.device)}

    def draw(self):
        """Random inputs of one level: resampling offset, proposal normals, acceptance uniforms."""
        n, d, s = self.num_particles, self.prior.dim, self.num_mcmc_steps
        return {"u": torch.rand((), generator=self.gen, device=self.device, dtype=torch.float64),
                "z": torch.randn((s, n, d), generator=self.gen, device=self.device, dtype=torch.float64),
                "a": torch.rand((s, n), generator=self.gen, device=self.device, dtype=torch.float64)}

    def level(self, state, rand):
        """One tempering level (reweight, evidence, resample, refit, mutate) in place on state."""
        x, logl, logprior, beta = state["x"], state["logl"], state["logprior"], state["beta"]
        delta = tempering_increment(logl, beta, self.target_ess)
        beta_new = beta + delta
        lw = torch.nan_to_num(delta * logl)
        state["log_evidence"].add_(torch.logsumexp(lw, 0) - math.log(self.num_particles))
        idx = systematic_resample(lw, rand["u"])
        x = x[idx]
        logl = logl[idx]
        logprior = logprior[idx]
        mean, _, chol = fit_gaussian(x, self.reg)
        logq = gaussian_logpdf_chol(x, mean, chol)
        acc = torch.zeros((), dtype=torch.float64, device=x.device)
        for s in range(self.num_mcmc_steps):
            xp = gaussian_from_normals(mean, chol, rand["z"][s])
            lpp = self.prior.log_prob(xp)
            llp = self._loglik(xp)
            lqp = gaussian_logpdf_chol(xp, mean, chol)
            cur = logprior + _tempered(beta_new, logl)
            prop = lpp + _tempered(beta_new, llp)
            log_alpha = (prop - cur) + (logq - lqp)
            take = (torch.log(rand["a"][s]) < log_alpha) & torch.isfinite(prop)
            x = torch.where(take[:, None], xp, x)
            logl = torch.where(take, llp, logl)
            logprior = torch.where(take, lpp, logprior)
            logq = torch.where(take, lqp, logq)
            acc = acc + take.to(torch.float64).mean()
        state["x"].copy_(x)
        state["logl"].copy_(logl)
        state["logprior"].copy_(logprior)
        state["beta"].copy_(beta_new)
        state["accept"].copy_(acc / self.num_mcmc_steps)
        state["delta"].copy_(delta)
        self.n_like_evals += self.num_particles * self.num_mcmc_steps
        return state

    def _physical(self, x):
        """Physical sampled parameters of the particles (the particles themselves here)."""
        return x

    def run(self, max_levels=1000, callback=None):
        """Levels until beta = 1; returns the reference's result fields (without file output)."""
        start = time.perf_counter()
        state = self.init_state()
        levels = []
        for _ in range(max_levels):
            self.level(state, self.draw())
            rec = {"beta": float(state["beta"]), "delta": float(state["delta"]),
                   "acceptance": float(state["accept"]), "log_evidence": float(state["log_evidence"])}
            levels.append(rec)
            if callback is not None:
                callback(rec)
            if rec["beta"] >= 1.0:
                break
        rows = self.prior.rows(self._physical(state["x"]))
        n = self.num_particles
        return {"posteriors": {k: rows[:, j] for j, k in enumerate(self.prior.names)},
                "log_likelihoods": state["logl"], "log_evidence": float(state["log_evidence"]),
                "weights": torch.full((n,), 1.0 / n, dtype=torch.float64, device=self.device),
                "levels": levels, "running_time": time.perf_counter() - start,
                "n_likelihood_evaluations": self.n_like_evals}
