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
                        d["mask"], df=d["df"], weight=d["weight"])
    return nc


def _minus_hessian(nc, row0, h):
    """-d^2 lnL / d row_i d row_j at row0 from central differences of the adjoint gradient,
    Richardson-combined over steps h and h / 2 (per coordinate)."""
    n = row0.size
    rows = []
    for k in range(n):
        for s in (h[k], -h[k], 0.5 * h[k], -0.5 * h[k]):
            r = row0.copy()
            r[k] += s
            rows.append(r)
    _, g = nc.lnl(np.array(rows), grad=True)
    g = g.cpu().numpy().reshape(n, 4, n)
    d1 = (g[:, 0] - g[:, 1]) / (2.0 * h[:, None])
    d2 = (g[:, 2] - g[:, 3]) / h[:, None]
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
