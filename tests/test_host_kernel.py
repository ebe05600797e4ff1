"""The fused kernel's per-(bin, row) device functions compiled for the host, under AddressSanitizer
and UndefinedBehaviorSanitizer, against the oracle.

response_row (the forward response the fused kernel and pes_response use), bin_prepare, load_row,
tdi_consts and mode_adjoint are compiled with a host stand-in for the HIP header
(tests/host_kernel/stub: the FP64 reciprocal becomes 1/x) twice: -O1 with ASan + UBSan (every
memory access and arithmetic operation checked) and -O3 without.  Both must give the oracle's
A/E/T to 1e-10 on every bin, including bins where all three link pairs take the reciprocal branch
of the sinc (|pi f L (1 +- k.n_a)| >= 0.01 for every pair), the case a spilling build of the
response kernel once returned wrong values for on the GPU.

interval_off and interval_on, the bounds by which the kernels find tiles that a duty cycle's off
periods or plateaus cover, are checked against brute force the same way
(tests/host_kernel/host_duty_off.cpp).
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import detector as odet
from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import inj_row, make_detector, particle_cloud, prior_rows

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pespace_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "host_kernel", "host_response.cpp")
STUB = os.path.join(ROOT, "tests", "host_kernel", "stub")
INTERP = (1e-4, 0.05, 1024)
MODE_IDS = {(2, 2): 0, (2, 1): 1, (3, 3): 2, (3, 2): 3, (4, 4): 4, (4, 3): 5}
PC_STRIDE, PC_MODE, PC_FCUT = 32 + 6 * 9 + 2, 32, 32 + 6 * 9
PC_SI = 3.085677581491367e16
CXX = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") else shutil.which("clang++")


def _pconst(row, modes):
    """The per-row constant block of k_setup (extrinsic part, FP64) from the oracle's pieces."""
    p = olk.as_dict(row)
    pc = np.zeros(PC_STRIDE)
    pc[0], pc[1], pc[2], pc[3] = p["coalescence_time"], 1.0 / p["luminosity_distance"], 1.0, p["luminosity_distance"]
    pv, qv, kv = odet.polarization_basis(p["ecliptic_longitude"], p["ecliptic_latitude"], p["polarization"])
    pc[4:7], pc[7:10], pc[10:13] = pv, qv, kv
    hm = oph.PhenomHM(p["chirp_mass"], p["mass_ratio"], p["chi_1"], p["chi_2"], 0.0, modes)
    dist = p["luminosity_distance"] * 1e6 * PC_SI
    fcut = 0.0
    for mi, md in enumerate(modes):
        kp, kc = oph.harmonic_factors(md, p["inclination"], p["reference_phase"])
        u = oph.mode_factor(md)
        Fp, Fc = kp * u / dist, 2.0 * kc * u / dist  # the cross factor is stored doubled
        cut = hm.mode_cut(md) / hm.Msec
        b = PC_MODE + 9 * mi
        pc[b:b + 5] = [Fp.real, Fp.imag, Fc.real, Fc.imag, cut]
        fcut = max(fcut, cut)
    pc[PC_FCUT] = fcut
    return pc


def _nodes(row, modes):
    p = olk.as_dict(row)
    hm = oph.PhenomHM(p["chirp_mass"], p["mass_ratio"], p["chi_1"], p["chi_2"], 0.0, modes)
    u0, du = oph.node_grid(*INTERP)
    t = np.zeros((len(modes), INTERP[2] + 1, 4))  # plus the zero padding node
    for mi, md in enumerate(modes):
        t[mi, :-1] = np.stack(oph.node_values(hm, md, u0, du, INTERP[2]), axis=-1)
    return t


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    if CXX is None:
        pytest.skip("no clang++")
    if not os.path.exists(os.path.join(CSRC, "trig_table.inc")):
        from pespace_amd import build as pbuild
        pbuild.gen_trig_inc()
        pbuild.gen_qnm_inc()
    d = tmp_path_factory.mktemp("host_kernel")
    out = {}
    base = [CXX, "-std=c++17", "-g", "-I", STUB, "-I", CSRC, "-Wno-unknown-attributes", "-o"]
    for name, flags in (("san", ["-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                                 "-fno-omit-frame-pointer"]), ("opt", ["-O3"])):
        exe = str(d / f"host_response_{name}")
        subprocess.run(base + [exe, DRIVER] + flags, check=True, capture_output=True, text=True)
        out[name] = exe
    return out, d


def _run(exe, d, det, labels, rows, modes):
    kind, arm, rot0, rev0 = det.orbit[:4]
    geo = tuple(det.orbit[4:]) or (odet.GEO_OMEGA_ROT, odet.GEO_LAMBDA_REF, odet.GEO_BETA_REF)
    u0, du = oph.node_grid(*INTERP)
    x = (np.log(det.f) - u0) / du
    j = np.clip(np.floor(x), 0, INTERP[2] - 2).astype(np.int32)
    s = x - j
    hdr = np.zeros(13, np.int32)
    hdr[:7] = [0 if kind == "helio" else 1, 15 if det.tdi_generation == "1.5" else 20, int(det.orthogonal),
               det.f.size, len(modes), INTERP[2] + 1, rows.shape[0]]
    hdr[7:7 + len(modes)] = [MODE_IDS[m] for m in modes]
    inp, outp = str(d / "in.bin"), str(d / "out.bin")
    with open(inp, "wb") as fh:
        hdr.tofile(fh)
        np.array([arm, rot0, rev0, *geo, du], dtype=np.float64).tofile(fh)
        det.f.astype(np.float64).tofile(fh)
        s.tofile(fh)
        j.tofile(fh)
        np.stack([_pconst(r, modes) for r in rows]).tofile(fh)
        np.stack([_nodes(r, modes) for r in rows]).tofile(fh)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, inp, outp], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    o = np.fromfile(outp).reshape(rows.shape[0], 3, det.f.size, 2)
    return o[..., 0] + 1j * o[..., 1]


def _all_pairs_reciprocal(det, row):
    """Bins where every link pair takes the reciprocal branch (both sinc arguments >= 0.01)."""
    p = olk.as_dict(row)
    _, _, k = odet.polarization_basis(p["ecliptic_longitude"], p["ecliptic_latitude"], p["polarization"])
    wf = oph.waveform_interpolated(p, det.f, oph.MODES, 0.0, *INTERP)["22"]
    kind, arm, rot0, rev0 = det.orbit[:4]
    n = odet.constellation(kind, arm, rot0, rev0, wf["tf"] + p["coalescence_time"], *det.orbit[4:])[:3]
    # (the (2,2) time above that mode's cut is not used by any kernel: those bins come out False)
    xL = np.pi * det.f * arm / 299792458.0
    ok = np.ones(det.f.size, bool)
    for na in n:
        kn = k @ na
        ok &= (np.abs(xL * (1 + kn)) >= 0.01) & (np.abs(xL * (1 - kn)) >= 0.01)
    return ok


CASES = [
    dict(),
    dict(orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3),
    dict(orbit=("geo", 1.7e8, 0.4, 1.1, 2 * np.pi / (5 * 86400.0), 0.7, 0.3), psd_model="Tianqin_GWSpace", fmin=1e-3),
    dict(gen="2.0", orthogonal=False),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_response_row_under_sanitizers_matches_oracle(binaries, case):
    exes, d = binaries
    modes = oph.MODES
    det, labels = make_detector(modes=modes, T=2 ** 13 * 5.0, **CASES[case])
    rows = np.vstack([inj_row(), particle_cloud(1, seed=3), prior_rows(1, seed=8)])
    got = {name: _run(exe, d, det, labels, rows, modes) for name, exe in exes.items()}
    np.testing.assert_array_equal(got["san"].shape, got["opt"].shape)
    for p in range(rows.shape[0]):
        ref = olk.response(olk.as_dict(rows[p]), det, modes, 0.0, INTERP)
        mask = _all_pairs_reciprocal(det, rows[p])
        if p < 2 and det.orbit[1] > 1e9:  # LISA arms near the injection: most bins take the
            assert mask.sum() > det.f.size // 2  # reciprocal branch on every pair (TianQin's short
            # arms, pi f L <= 0.09, mostly take the series branch: the other case)
        for ci, c in enumerate(labels):
            r = np.asarray(ref[c])
            tol = 1e-10
            if c == "T":  # the oracle's (X+Y+Z)/sqrt3 cancels to ~ (pi f L)^2 |X|
                xyz = np.sqrt(np.sum(np.abs(ref[labels[0]]) ** 2 + np.abs(ref[labels[1]]) ** 2))
                tol = max(tol, 64 * np.finfo(float).eps * xyz / np.linalg.norm(r))
            for name in ("san", "opt"):
                g = got[name][p, ci]
                assert np.linalg.norm(g - r) <= tol * np.linalg.norm(r), (name, p, c)
                assert np.linalg.norm((g - r)[mask]) <= tol * np.linalg.norm(r[mask]), (name, p, c, "reciprocal bins")


def test_duty_interval_bounds_are_sound(tmp_path):
    """interval_off and interval_on (whether a node interval's t_f + t_c lies in an off period, or
    on a plateau clear of the ramps, of the window: k_duty_off, k_duty_items) against brute force
    over 60,000 random intervals near the edges and ramp ends of three windows (host build under
    ASan/UBSan): every marked interval sampled at 4001 points with the fused kernel's own t and
    window has w = 0 (w = 1 and w' = 0); weakened bounds (endpoints only, no margin) fail the
    same sampling on some intervals, so the sampling can tell."""
    if CXX is None:
        pytest.skip("no clang++")
    if not os.path.exists(os.path.join(CSRC, "trig_table.inc")):
        from pespace_amd import build as pbuild
        pbuild.gen_trig_inc()
        pbuild.gen_qnm_inc()
    exe = str(tmp_path / "host_duty_off")
    src = os.path.join(ROOT, "tests", "host_kernel", "host_duty_off.cpp")
    subprocess.run([CXX, "-std=c++17", "-g", "-I", STUB, "-I", CSRC, "-Wno-unknown-attributes", "-O1",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                    "-o", exe, src], check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    counts = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
    assert counts["marked"] > 500 and counts["marked_on"] > 500, counts
    assert counts["weak_misses"] > 0 and counts["weak_on_misses"] > 0, counts
