"""Write the golden fixtures of the parity tests with the numpy FP64 oracle.

    python -m tests.golden.make_golden          (from the repository root)

Each case is a detector (orbit, PSD, TDI generation, grid, noise seed), a mode set, an
interpolation grid and six parameter rows near the canonical injection
(examples/pe_scripts/full_pe_ti_lisa_xas.py:24-52); the data are the seeded noise realization
alone (the noise generator is bit-identical between the oracle and the product), so the
inputs are fully defined by the stored specification.  Stored expected outputs: lnL of every
row, and for row 0 the A/E/T (or X/Y/Z) response and the waveform container {plus, cross,
tf} of every mode at 24 sample bins.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import detector as odet  # noqa: E402
from oracle import likelihood as olk  # noqa: E402
from oracle import phenom as oph  # noqa: E402
from tests._setup import inj_row, make_detector, particle_cloud  # noqa: E402

PATH = os.path.join(HERE, "golden_parity.npz")
HM = [list(m) for m in oph.MODES]

CASES = {
    "lisa_22_tdi15": dict(orbit="LISA_analytic", psd_model="LISA_SciRDv1", gen="1.5", modes=[[2, 2]]),
    "lisa_hm_tdi20": dict(orbit="LISA_analytic", psd_model="LISA_SciRDv1", gen="2.0", modes=HM),
    "taiji_hm": dict(orbit="Taiji_analytic", psd_model="Taiji_TDC", gen="1.5", modes=HM),
    "tianqin_hm": dict(orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", gen="1.5", modes=HM, fmin=1e-3),
    "lisa_22_xyz": dict(orbit="LISA_analytic", psd_model="LISA_SciRDv1", gen="1.5", modes=[[2, 2]],
                        orthogonal=False),
}
COMMON = dict(T=131072.0 * 5, dt=5.0, fmin=1e-4, fmax=0.05, seed=17, interp=[1e-4, 0.05, 1024])


def case_spec(name):
    spec = dict(COMMON)
    spec.update(CASES[name])
    return spec


def compute(spec):
    modes = tuple(tuple(m) for m in spec["modes"])
    interp = tuple(spec["interp"])
    d, labels = make_detector(orbit=spec["orbit"], psd_model=spec["psd_model"], gen=spec["gen"], T=spec["T"],
                              dt=spec["dt"], fmin=spec["fmin"], fmax=spec["fmax"], seed=spec["seed"], modes=modes,
                              interp=interp, orthogonal=spec.get("orthogonal", True), inject=False)
    rows = np.vstack([inj_row(), particle_cloud(5, seed=31, rel=1e-4)])
    lnl = olk.lnl_batch(rows, [d], modes, 0.0, interp)
    p0 = olk.as_dict(rows[0])
    resp = olk.response(p0, d, modes, 0.0, interp)
    idx = np.unique(np.linspace(0, d.f.size - 1, 24).astype(np.int64))
    wf = olk.waveform(p0, d.f[idx], modes, 0.0, interp)
    wfs = np.stack([np.stack([wf[f"{l}{m}"]["plus"], wf[f"{l}{m}"]["cross"],
                              wf[f"{l}{m}"]["tf"].astype(np.complex128)]) for l, m in modes])
    return dict(rows=rows, lnl=lnl, idx=idx, resp=np.stack([resp[c][idx] for c in labels]), wf=wfs,
                labels=np.array(labels))


def main():
    out = {}
    for name in CASES:
        spec = case_spec(name)
        r = compute(spec)
        out[f"{name}/spec"] = np.array(json.dumps(spec))
        for k, v in r.items():
            out[f"{name}/{k}"] = v
    np.savez_compressed(PATH, **out)
    print("wrote", PATH, os.path.getsize(PATH), "bytes")


if __name__ == "__main__":
    main()
