"""Fisher-matrix helpers on the CPU: sampling from N(0, F^-1) (fisher.py:165-232) on a
mixed-unit matrix, the angle wrapping (fisher.py:235-311), and the channel check of
fisher_matrix (fisher.py:82-87), which fails before any GPU work."""

from collections import namedtuple

import numpy as np
import pytest

from pespace_amd.inference.fisher import fisher_matrix, sample_from_fisher, wrap_angle_params


def _mixed_unit_fisher(n=6, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)  # well conditioned in natural units
    D = np.diag(10.0 ** np.linspace(-6, 6, n))
    return D @ A @ D, A, D


def test_sample_from_fisher_covariance_and_seed():
    F, A, D = _mixed_unit_fisher()
    assert np.linalg.cond(F) > 1e20
    x = sample_from_fisher(F, 200000, rkey_seed=3)
    assert x.shape == (200000, 6)
    y = x @ D  # natural units: covariance A^-1
    np.testing.assert_allclose(np.cov(y.T), np.linalg.inv(A), rtol=0.03, atol=3e-3 * np.abs(np.linalg.inv(A)).max())
    assert np.all(np.abs(y.mean(0)) < 5 * np.sqrt(np.diag(np.linalg.inv(A)) / 200000))
    np.testing.assert_array_equal(sample_from_fisher(F, 1000, rkey_seed=3), sample_from_fisher(F, 1000, rkey_seed=3))
    assert not np.array_equal(sample_from_fisher(F, 1000, rkey_seed=4), sample_from_fisher(F, 1000, rkey_seed=3))


def test_wrap_angle_params_ranges_and_identities():
    rng = np.random.default_rng(1)
    v = rng.uniform(-20, 20, 1000)
    s = {"ecliptic_longitude": v, "ecliptic_latitude": v, "inclination": v, "polarization": v,
         "reference_phase": v, "chirp_mass": v}
    w = wrap_angle_params(s)
    assert w is not s and s["ecliptic_longitude"] is v and w["chirp_mass"] is v
    assert np.all((w["ecliptic_longitude"] >= 0) & (w["ecliptic_longitude"] < 2 * np.pi))
    assert np.all((w["polarization"] >= 0) & (w["polarization"] < np.pi))
    assert np.all(np.abs(w["ecliptic_latitude"]) <= np.pi / 2)
    assert np.all((w["inclination"] >= 0) & (w["inclination"] <= np.pi))
    # periodic angles: the same point; bounded ones: a reflection (same sine of latitude /
    # cosine of inclination up to the fold)
    np.testing.assert_allclose(np.cos(w["ecliptic_longitude"]), np.cos(v), atol=1e-12)
    np.testing.assert_allclose(np.sin(2 * w["polarization"]), np.sin(2 * v), atol=1e-12)
    np.testing.assert_allclose(np.sin(w["ecliptic_latitude"]), np.sin(v), atol=1e-12)
    np.testing.assert_allclose(np.cos(w["inclination"]), np.cos(v), atol=1e-12)
    # in-range values are unchanged
    inside = {"ecliptic_latitude": np.array([-1.2, 0.3]), "inclination": np.array([0.1, 3.0])}
    w2 = wrap_angle_params(inside)
    np.testing.assert_allclose(w2["ecliptic_latitude"], inside["ecliptic_latitude"], rtol=1e-15)
    np.testing.assert_allclose(w2["inclination"], inside["inclination"], rtol=1e-15)
    # reflections at the boundaries
    eps = 1e-3
    np.testing.assert_allclose(wrap_angle_params({"ecliptic_latitude": np.pi / 2 + eps})["ecliptic_latitude"],
                               np.pi / 2 - eps, rtol=1e-12)
    np.testing.assert_allclose(wrap_angle_params({"inclination": -eps})["inclination"], eps, rtol=1e-9)
    with pytest.raises(TypeError, match="mapping"):
        wrap_angle_params([1.0, 2.0])


def test_fisher_matrix_requires_orthogonal_channels():
    class _Lk:
        channels = ("X", "Y")
        likelihood_param_type = namedtuple("P", ("a",))

    with pytest.raises(ValueError, match="orthogonal channels"):
        fisher_matrix(_Lk(), {"a": 1.0})
