"""Fisher matrix from the adjoint kernel (minus the Hessian of the noiseless lnL) against the
numpy oracle's Fisher from finite differences of its TDI response, F_ij = 4 df Re sum
conj(ds_i) ds_j / S (fisher.py:27-162); distance scaling and batches."""

from collections import namedtuple

import numpy as np
import pytest

from oracle import likelihood as olk
from tests._setup import INJ, add_to_core, inj_row, make_detector

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)
MODES = ((2, 2),)


class _Lk:
    """The parts of WhittleLikelihoodHip that fisher_matrix uses, over a test context."""

    channels = ("A", "E", "T")
    likelihood_param_type = namedtuple("P", olk.PARAM_NAMES)

    def __init__(self, core):
        self.core = core

    def row(self, p):
        return np.array([float(p[k]) for k in olk.PARAM_NAMES])


def _setup():
    from pespace_amd.core import LikelihoodCore
    d, labels = make_detector(modes=MODES, T=131072.0 * 5, dt=5.0, seed=21)
    core = LikelihoodCore(modes=["22"], n_nodes=INTERP[2], f_node_min=INTERP[0], f_node_max=INTERP[1])
    add_to_core(core, d, labels)
    return d, _Lk(core)


# central-difference steps of the oracle response (well inside the linear regime)
STEPS = np.array([1e-6 * INJ["chirp_mass"], 1e-6, 1e-6, 1e-6, 1e-4 * INJ["luminosity_distance"], 1e-5, 1e-5,
                  1e-6, 1e-6, 1e-6, 1e-3])


def _oracle_fisher(d, row):
    def resp(r):
        out = olk.response(olk.as_dict(r), d, MODES, 0.0, INTERP)
        return np.stack([out[c] for c in ("A", "E", "T")])

    w = np.stack([2.0 * d.df / d.psd[c] for c in ("A", "E", "T")])
    ds = []
    for k in range(row.size):
        def at(s):
            r = row.copy()
            r[k] += s
            return resp(r)
        h = STEPS[k]
        d1 = (at(h) - at(-h)) / (2 * h)
        d2 = (at(h / 2) - at(-h / 2)) / h
        ds.append((4 * d2 - d1) / 3)
    ds = np.array(ds)  # (n, 3, N)
    return 2.0 * np.einsum("icn,jcn,cn->ij", ds.conj(), ds, w).real


def _normalised(F):
    s = np.sqrt(np.diag(F))
    return F / np.outer(s, s)


def test_fisher_matches_oracle_response_differences():
    from pespace_amd.inference.fisher import fisher_matrix
    d, lk = _setup()
    F = fisher_matrix(lk, INJ)
    Fo = _oracle_fisher(d, inj_row())
    print("diag rel", np.diag(F) / np.diag(Fo) - 1)
    np.testing.assert_allclose(np.diag(F), np.diag(Fo), rtol=2e-5)
    np.testing.assert_allclose(_normalised(F), _normalised(Fo), atol=2e-5)
    np.testing.assert_array_equal(F, F.T)
    assert np.all(np.linalg.eigvalsh(_normalised(F)) > 0)


def test_fisher_distance_scaling_and_batch():
    from pespace_amd.inference.fisher import fisher_matrix
    _, lk = _setup()
    F1 = fisher_matrix(lk, INJ)
    p2 = dict(INJ)
    p2["luminosity_distance"] = 2.0 * INJ["luminosity_distance"]
    F2 = fisher_matrix(lk, p2)
    k = olk.PARAM_NAMES.index("luminosity_distance")
    others = [i for i in range(11) if i != k]
    # s ~ 1/d_L: the other block scales as 1/d_L^2, the mixed row as 1/d_L^3, F_dd as 1/d_L^4
    np.testing.assert_allclose(F2[np.ix_(others, others)], F1[np.ix_(others, others)] / 4, rtol=1e-5,
                               atol=1e-6 * np.abs(F1[np.ix_(others, others)]).max())
    np.testing.assert_allclose(F2[k, k], F1[k, k] / 16, rtol=1e-5)
    batch = {key: np.array([v, p2[key]]) for key, v in INJ.items()}
    Fb = fisher_matrix(lk, batch)
    assert Fb.shape == (2, 11, 11)
    np.testing.assert_allclose(Fb[0], F1, rtol=1e-12)
    np.testing.assert_allclose(Fb[1], F2, rtol=1e-12)


def test_fisher_at_the_mass_ratio_edge():
    """An equal-mass fiducial point (q = 1, the edge of the mass-ratio domain) takes a one-sided
    stencil along q; the matrix is finite, positive definite and continuous with the one just
    inside the domain (central stencils)."""
    from pespace_amd.inference.fisher import fisher_matrix
    _, lk = _setup()
    p1 = dict(INJ, mass_ratio=1.0)
    p0 = dict(INJ, mass_ratio=1.0 - 1e-5)
    F1 = fisher_matrix(lk, p1)
    F0 = fisher_matrix(lk, p0)
    assert np.all(np.isfinite(F1))
    assert np.all(np.linalg.eigvalsh(_normalised(F1)) > 0)
    np.testing.assert_allclose(np.diag(F1), np.diag(F0), rtol=1e-3)
    np.testing.assert_allclose(_normalised(F1), _normalised(F0), atol=1e-3)
