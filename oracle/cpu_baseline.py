"""Driver of the oracle's C++/OpenMP restatement (``oracle/cpu_lnl.cpp``): the CPU baseline.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): bench.py's cpu_baseline leg and tests/
use it; the product never does.

The waveform's node tables (A, dA/du, Psi, dPsi/du of every mode on the log nodes) come from
the numpy oracle (``oracle/phenom.py``); for the gradient, their derivatives in chirp mass, mass
ratio and the two spins are central differences of those tables (relative step 1e-6, accurate
to ~1e-9 relative), and the harmonic factors' inclination derivatives likewise.  Everything per
bin (interpolation, t_c shift, orbit, links, TDI, Whittle sum and, for the gradient, its
forward-mode derivatives) runs in C++ with OpenMP over the bins.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import phenom as ph
from .build_cpu import LIB

_lib = None
c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int)
PARAM_NAMES = ("chirp_mass", "mass_ratio", "chi_1", "chi_2", "luminosity_distance", "reference_phase",
               "inclination", "ecliptic_longitude", "ecliptic_latitude", "polarization", "coalescence_time")
C_SI = 299792458.0


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (python oracle/build_cpu.py)")
        L = ctypes.CDLL(LIB)
        L.cpu_lnl_row.restype = ctypes.c_int
        L.cpu_lnl_row.argtypes = [ctypes.c_int, c_ip, c_dp, c_ip, c_ip, ctypes.c_int, c_dp, c_dp, c_dp, c_dp,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, c_dp, c_dp, c_dp,
                                  c_dp, c_dp, c_dp, ctypes.c_int, ctypes.c_int, c_dp, c_dp]
        _lib = L
    return _lib


def _p(a, t=c_dp):
    return a.ctypes.data_as(t) if a is not None else None


class Network:
    """Concatenated detector arrays.  dets: objects with f, data {c: complex}, psd {c: real},
    df (scalar or per-bin), orbit (kind, arm[m], rot0, rev0[, omega_rot, lambda_ref, beta_ref]),
    tdi_generation ("1.5"/"2.0"), channels, orthogonal and optionally duty (period, on_duration,
    start_time, taper) (the oracle's likelihood.Detector)."""

    def __init__(self, dets):
        from .detector import GEO_BETA_REF, GEO_LAMBDA_REF, GEO_OMEGA_ROT
        n = [d.f.size for d in dets]
        self.NB = int(sum(n))
        self.det_n = np.array(n, dtype=np.int32)
        self.det_off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int32)
        self.f = np.ascontiguousarray(np.concatenate([d.f for d in dets]), dtype=np.float64)
        self.w = np.zeros((3, self.NB))
        self.dre = np.zeros((3, self.NB))
        self.dim = np.zeros((3, self.NB))
        det_i, det_d = [], []
        for d, o, k in zip(dets, self.det_off, n):
            labels = ("A", "E", "T") if d.orthogonal else ("X", "Y", "Z")
            for c, lab in enumerate(labels):
                if lab in d.channels:
                    self.w[c, o:o + k] = 2.0 * np.asarray(d.df) / np.asarray(d.psd[lab])
                    self.dre[c, o:o + k] = np.real(d.data[lab])
                    self.dim[c, o:o + k] = np.imag(d.data[lab])
            kind, arm, rot0, rev0 = d.orbit[:4]
            geo = tuple(d.orbit[4:]) or (GEO_OMEGA_ROT, GEO_LAMBDA_REF, GEO_BETA_REF)
            duty = getattr(d, "duty", None)
            det_i += [0 if kind == "helio" else 1, 15 if d.tdi_generation == "1.5" else 20, int(d.orthogonal),
                      int(duty is not None)]
            det_d += [arm / C_SI, rot0, rev0, *geo, *(duty if duty is not None else (0.0, 0.0, 0.0, 0.0))]
        self.det_i = np.array(det_i, dtype=np.int32)
        self.det_d = np.array(det_d, dtype=np.float64)
        self.n_det = len(dets)


class RowInputs:
    """Per-row waveform inputs of the C++ code (node tables and harmonic factors)."""

    def __init__(self, row, modes, f_ref, interp, grad=False, rel=1e-6):
        p = dict(zip(PARAM_NAMES, (float(x) for x in row)))
        fmin, fmax, n_nodes = interp
        self.u0, self.du = ph.node_grid(fmin, fmax, n_nodes)
        self.n_nodes = n_nodes
        self.n_modes = len(modes)
        self.row = np.ascontiguousarray(row, dtype=np.float64)

        def tables(q):
            hm = ph.PhenomHM(q["chirp_mass"], q["mass_ratio"], q["chi_1"], q["chi_2"], f_ref, modes)
            t = np.stack([np.stack(ph.node_values(hm, md, self.u0, self.du, n_nodes), axis=-1) for md in modes])
            return hm, t  # (modes, nodes, 4)

        hm, self.nodes = tables(p)
        self.nodes = np.ascontiguousarray(self.nodes)
        self.mode_cut = np.array([hm.mode_cut(md) / hm.Msec for md in modes])
        self.nodes_grad = None
        if grad:
            g = np.zeros((self.n_modes, n_nodes, 4, 4))
            for i, key in enumerate(("chirp_mass", "mass_ratio", "chi_1", "chi_2")):
                h = rel * max(abs(p[key]), 1e-3)
                up, dn = dict(p), dict(p)
                up[key] += h
                dn[key] -= h
                if key == "mass_ratio" and up[key] > 1.0:  # one-sided at the edge q = 1
                    up[key], dn[key], scale = p[key], p[key] - h, h
                else:
                    scale = 2.0 * h
                g[:, :, i, :] = (tables(up)[1] - tables(dn)[1]) / scale
            self.nodes_grad = np.ascontiguousarray(g)

        def kfac(iota, phi):
            out = np.zeros((self.n_modes, 4))
            for m, md in enumerate(modes):
                kp, kc = ph.harmonic_factors(md, iota, phi)
                mf = ph.mode_factor(md)
                kp, kc = kp * mf, kc * mf
                out[m] = [kp.real, kp.imag, kc.real, kc.imag]
            return out

        self.kfac = np.ascontiguousarray(kfac(p["inclination"], p["reference_phase"]))
        self.kfac_d = None
        if grad:
            h = 1e-6
            d_iota = (kfac(p["inclination"] + h, p["reference_phase"])
                      - kfac(p["inclination"] - h, p["reference_phase"])) / (2 * h)
            # K ~ exp(-i m (pi/2 - phi_ref)): d/d phi_ref = i m K
            d_phi = np.zeros_like(self.kfac)
            for m, md in enumerate(modes):
                kp = self.kfac[m, 0] + 1j * self.kfac[m, 1]
                kc = self.kfac[m, 2] + 1j * self.kfac[m, 3]
                dkp, dkc = 1j * md[1] * kp, 1j * md[1] * kc
                d_phi[m] = [dkp.real, dkp.imag, dkc.real, dkc.imag]
            self.kfac_d = np.ascontiguousarray(np.stack([d_iota, d_phi], axis=1))


def lnl_row(net, ri, grad=False, stride=1, threads=1):
    """lnL (and the gradient) of one row on the network; every stride-th bin, weights x stride."""
    out = np.zeros(1)
    g = np.zeros(11) if grad else None
    rc = lib().cpu_lnl_row(net.n_det, _p(net.det_i, c_ip), _p(net.det_d), _p(net.det_off, c_ip),
                           _p(net.det_n, c_ip), net.NB, _p(net.f), _p(net.w), _p(net.dre), _p(net.dim), ri.n_modes,
                           ri.n_nodes, ri.u0, ri.du, _p(ri.nodes), _p(ri.nodes_grad), _p(ri.kfac), _p(ri.kfac_d),
                           _p(ri.mode_cut), _p(ri.row), int(stride), int(threads), _p(out), _p(g))
    if rc != 0:
        raise RuntimeError(f"cpu_lnl_row failed ({rc})")
    return (out[0], g) if grad else out[0]
