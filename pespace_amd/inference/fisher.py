"""Fisher matrix, Gaussian sampling from it and angle wrapping (pespace/inference/fisher.py)
on the hip backend.

fisher_matrix (fisher.py:27-162): F_ij = 4 df Re sum_{c,f} conj(ds_i) ds_j / S over the
detectors and channels, rows and columns in the order of ``likelihood_param_type``.  The
reference differentiates the response with forward-mode AD; here F is minus the Hessian of
the Whittle lnL for noiseless data equal to the model response at the fiducial point, the
identity the reference states (fisher.py:78-81): with d = s(theta0),
  d^2 lnL / d theta_i d theta_j (theta0) = -<ds_i, ds_j> - <d^2 s, s(theta0) - d> = -F_ij.
The Hessian comes from central differences of the exact adjoint gradient (the fused kernel's
lnL + gradient launch, all 4 n shifted rows in one batch), Richardson-combined over steps
h and h/2, with h_k = 1e-3 sigma_k where sigma_k = F_kk^{-1/2} from a pilot pass at
h_k = 1e-9 max(|theta_k|, 1): the truncation error is then ~(h/sigma)^4 relative and the
rounding error of the gradient differences a few ulp of F.  Parametrizations other than the
core's (chirp mass, mass ratio, ...) are handled by the chain rule F = J^T F_core J with J the
Jacobian of the map to the core's row (exact at theta0, where the noiseless gradient
vanishes).

sample_from_fisher (fisher.py:165-232): zero-mean samples with covariance F^{-1} through the
unit-diagonal scaling F' = F / (s s^T), s = sqrt(diag F), the Cholesky factor F' = L L^T and
the triangular solve L^T x' = z, x = x' / s.  The normal variates come from torch's generator
seeded with ``rkey_seed`` (not JAX's stream).

wrap_angle_params (fisher.py:235-311): longitude mod 2 pi, latitude and inclination folded by
reflection, polarization mod pi, reference phase mod 2 pi.
"""

from __future__ import annotations

import numpy as np
import torch

from ..core import PARAM_NAMES, LikelihoodCore

_PILOT = 1e-9
_STEP = 1e-3


def _noiseless_core(core, row):
    """A context with core's model and detectors whose data are the response at row."""
    nc = LikelihoodCore(modes=core.modes, reference_frequency=core.reference_frequency, n_nodes=core.n_nodes,
                        f_node_min=core.f_node_min, f_node_max=core.f_node_max, device=core.device)
    r = torch.as_tensor(np.asarray(row, dtype=np.float64)[None, :])
    for i, d in enumerate(core._dets):
        resp = core.response(i, r)[0].cpu().numpy()
        data = np.where(np.asarray(d["mask"])[:, None] != 0, resp, 0.0)
        nc.add_detector(d["kind"], d["arm"], d["rot0"], d["rev0"], d["gen"], d["orth"], d["f"], data, d["psd"],
                        d["mask"], df=d["df"], weight=d["weight"], omega_rotation=d["omega"],
                        lambda_ref=d["lam"], beta_ref=d["bet"], duty_cycle=d["duty"])
    # the source context's schedule (autotuned or set by the caller) instead of the default; the
    # rows per work item stay chosen per batch (the Fisher batches hold 45 rows)
    nc.set_schedule(*core.schedule())
    return nc


def _in_domain(row):
    """The rows the kernel evaluates (the others give lnL = -inf): chirp mass and distance
    positive, mass ratio in (0, 1], |chi| < 1."""
    return row[0] > 0.0 and 0.0 < row[1] <= 1.0 and abs(row[2]) < 1.0 and abs(row[3]) < 1.0 and row[4] > 0.0


def _minus_hessian(nc, row0, h):
    """-d^2 lnL / d row_i d row_j at row0 from differences of the adjoint gradient,
    Richardson-combined over steps h and h / 2 (per coordinate).

    Central differences where row0 +- h stay in the domain; at its edge (e.g. an equal-mass
    fiducial point, q = 1) the one-sided second-order stencil
    D(h) = s (-3 g(0) + 4 g(s h) - g(2 s h)) / (2 h) pointing inwards (s = +-1)."""
    n = row0.size
    side = np.zeros(n)  # 0: central, +1 / -1: one-sided towards +h / -h
    rows = [row0.copy()]
    for k in range(n):
        up, dn = row0.copy(), row0.copy()
        up[k] += h[k]
        dn[k] -= 2.0 * h[k]
        ok_up, ok_dn = _in_domain(up), _in_domain(dn)
        if not ok_up:
            side[k] = -1.0
        elif not ok_dn:
            up2 = row0.copy()
            up2[k] += 2.0 * h[k]
            side[k] = 1.0 if _in_domain(up2) else 0.0
        if side[k] == 0.0:
            steps = (h[k], -h[k], 0.5 * h[k], -0.5 * h[k])
        else:
            sh = side[k] * h[k]
            steps = (sh, 2.0 * sh, 0.5 * sh, sh)
        for st in steps:
            r = row0.copy()
            r[k] += st
            if not _in_domain(r):
                raise ValueError(f"Fisher matrix: the fiducial {PARAM_NAMES[k]} = {row0[k]!r} lies too close to "
                                 f"both ends of its domain for a step of {h[k]:.3g}")
            rows.append(r)
    lnl, g = nc.lnl(np.array(rows), grad=True)
    lnl = lnl.cpu().numpy()
    g = g.cpu().numpy()
    if not np.all(np.isfinite(lnl)):
        bad = sorted({PARAM_NAMES[(i - 1) // 4] for i in np.nonzero(~np.isfinite(lnl))[0] if i > 0})
        raise ValueError(f"Fisher matrix: lnL is not finite at the fiducial point or its shifts along {bad}")
    g0 = g[0]
    g = g[1:].reshape(n, 4, n)
    hk = h[:, None]
    cen = side == 0.0
    d1 = np.where(cen[:, None], (g[:, 0] - g[:, 1]) / (2.0 * hk),
                  side[:, None] * (-3.0 * g0[None, :] + 4.0 * g[:, 0] - g[:, 1]) / (2.0 * hk))
    d2 = np.where(cen[:, None], (g[:, 2] - g[:, 3]) / hk,
                  side[:, None] * (-3.0 * g0[None, :] + 4.0 * g[:, 2] - g[:, 3]) / hk)
    H = (4.0 * d2 - d1) / 3.0  # H[k, :] = d grad / d row_k
    F = -H
    return 0.5 * (F + F.T)


def _fisher_core(nc, row0):
    h = _PILOT * np.maximum(np.abs(row0), 1.0)
    F = _minus_hessian(nc, row0, h)
    diag = np.diag(F)
    good = np.isfinite(diag) & (diag > 0.0)
    h = np.where(good, _STEP / np.sqrt(np.where(good, diag, 1.0)), h)
    return _minus_hessian(nc, row0, h)


def _to_row_jacobian(likelihood, names, theta0):
    """d row / d theta (n_core x n_fields): a permutation for the core's own parameter names,
    otherwise central differences of likelihood.row (a smooth map of the masses)."""
    def to_row(t):
        return likelihood.row(dict(zip(names, t)))

    row0 = to_row(theta0)
    if sorted(names) == sorted(PARAM_NAMES):
        J = np.zeros((row0.size, len(names)))
        for j, k in enumerate(names):
            J[PARAM_NAMES.index(k), j] = 1.0
        return row0, J
    J = np.zeros((row0.size, len(names)))
    for j in range(len(names)):
        hj = 1e-6 * max(abs(theta0[j]), 1.0)
        tp, tm = theta0.copy(), theta0.copy()
        tp[j] += hj
        tm[j] -= hj
        J[:, j] = (to_row(tp) - to_row(tm)) / (2.0 * hj)
    J[np.abs(J) < 1e-12 * np.abs(J).max()] = 0.0
    return row0, J


def fisher_matrix(likelihood, fiducial_params, batch_size=None):
    """Fisher matrix at one or more fiducial parameter sets (fisher.py:27-162).

    likelihood: WhittleLikelihoodHip.  fiducial_params: dict with every field of
    ``likelihood.likelihood_param_type`` (scalars, or 1-D arrays of a common length for a
    batch, scalars broadcast).  Returns an (n, n) array, or (B, n, n) for a batch; the
    observed data of the detectors are not used.  ``batch_size`` is accepted for the
    reference's signature; the sets are evaluated one after another.
    """
    if not set(likelihood.channels) <= {"A", "E", "T"}:
        raise ValueError("fisher_matrix requires the independent orthogonal channels ('A', 'E', 'T'), but got the "
                         f"channels {tuple(likelihood.channels)}.")
    names = tuple(likelihood.likelihood_param_type._fields)
    vals = {k: np.asarray(fiducial_params[k], dtype=np.float64) for k in names}
    shape = np.broadcast_shapes(*(v.shape for v in vals.values()))
    if len(shape) > 1:
        raise ValueError("fiducial parameters must be scalars or 1-D arrays")
    sets = [{k: float(v) for k, v in vals.items()}] if shape == () else \
        [{k: float(np.broadcast_to(v, shape)[i]) for k, v in vals.items()} for i in range(shape[0])]
    out = []
    for p in sets:
        theta0 = np.array([p[k] for k in names])
        row0, J = _to_row_jacobian(likelihood, names, theta0)
        nc = _noiseless_core(likelihood.core, row0)
        try:
            Fc = _fisher_core(nc, row0)
        finally:
            nc.close()
        F = J.T @ Fc @ J
        out.append(0.5 * (F + F.T))
    return out[0] if shape == () else np.stack(out)


def sample_from_fisher(fisher, num, rkey_seed=0):
    """Zero-mean samples with covariance fisher^{-1} (fisher.py:165-232); (num, N) array."""
    F = torch.as_tensor(np.asarray(fisher, dtype=np.float64))
    s = torch.sqrt(torch.diagonal(F))
    Fs = F / torch.outer(s, s)
    L = torch.linalg.cholesky(Fs)
    g = torch.Generator().manual_seed(int(rkey_seed))
    z = torch.randn((int(num), F.shape[0]), generator=g, dtype=torch.float64)
    x = torch.linalg.solve_triangular(L.T, z.T, upper=True).T
    return (x / s).numpy()


def wrap_angle_params(samples):
    """Angles into their canonical ranges (fisher.py:235-311); returns a shallow copy."""
    try:
        keys = set(samples.keys())
    except AttributeError as error:
        raise TypeError("wrap_angle_params expects a mapping of parameter names to values (e.g. a dict or a pandas "
                        f"DataFrame), but got {type(samples).__name__}.") from error
    out = samples.copy()
    two_pi = 2.0 * np.pi
    if "ecliptic_longitude" in keys:
        out["ecliptic_longitude"] = out["ecliptic_longitude"] % two_pi
    if "ecliptic_latitude" in keys:
        lat = out["ecliptic_latitude"]
        out["ecliptic_latitude"] = np.pi / 2 - abs((lat + np.pi / 2) % two_pi - np.pi)
    if "inclination" in keys:
        inc = out["inclination"]
        out["inclination"] = np.pi - abs(inc % two_pi - np.pi)
    if "polarization" in keys:
        out["polarization"] = out["polarization"] % np.pi
    if "reference_phase" in keys:
        out["reference_phase"] = out["reference_phase"] % two_pi
    return out


__all__ = ["fisher_matrix", "sample_from_fisher", "wrap_angle_params", "PARAM_NAMES"]
