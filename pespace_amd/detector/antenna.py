"""Detector facade, the Marsat-Baker FD response model and the time-domain response, hip backend.

Mirrors pespace/detector/antenna.py:117-379: ``InterferometerAntenna`` owns the
TDI data, orbit, response model and TDI combination.  As in the reference,
``update_detector_response`` runs two stages on the GPU: the response model's
``update_single_link_response`` stores the six single links on the detector
(``single_link_response``, read back by ``single_link_response_array``), and the TDI
combination's ``update_tdi_response`` combines them into A/E/T (or X/Y/Z);
``inject_signal`` adds the result to the data.  The JAX backend's pure functions
``get_response_jit`` (six links) and ``get_combination_jit`` (three channels) are
provided too (_antenna_jax.py:82-195, _tdi_jax.py:70-102).  The waveform may be a
hip waveform model (``pespace_amd.waveform.IMRPhenomD/IMRPhenomHM``, evaluated on
the device) or any reference-style container ``{"plus","cross","tf"}`` / mode-keyed
dict of those (antenna.py:438-458).  The likelihood itself never materialises the
links: the fused kernel takes them straight into the residual.  ``TDResponseModelCornish2003``
with ``tdi.TDMichelsonConstantEqualArm`` is the time-domain counterpart (real time series on
the TDI model's extended sample grid).
"""

from __future__ import annotations

from collections import namedtuple

import numpy as np
import torch

from .orbit import OrbitModelBase, available_orbit_models

BACKENDS = ("hip",)
LINK_NAMES = ("link12", "link21", "link23", "link32", "link31", "link13")  # antenna.py:24-31

# An observation duty cycle (no counterpart in the reference): the detector observes for
# on_duration seconds of every period, from start_time + n period, with smoothstep ramps of
# taper seconds inside each window; each mode's response at a frequency is weighted by the
# window at that frequency's time t_f + t_c.
DutyCycle = namedtuple("DutyCycle", ["period", "on_duration", "start_time", "taper"], defaults=(0.0, 86400.0))


class FDResponseModelMarset2018:
    """Marsat et al. 2018 (arXiv:1806.10734) FD single-link response, fused with TDI on the device."""

    def __init__(self, backend="hip"):
        if backend not in BACKENDS:
            raise ValueError(f"Invalid backend: {backend}")
        self.backend = backend
        self.detector = None

    def init_single_link_response_model(self, detector):
        self.detector = detector
        detector.single_link_response = None

    def get_response_jit(self, waveform, lam, beta, psi, tc):
        """Six single links (link12, link21, link23, link32, link31, link13), each a complex (N,)
        tensor on the GPU (_antenna_jax.py:82-195); nothing is stored on the detector."""
        links = self.detector._links(waveform, lam, beta, psi, tc)
        return tuple(links[i] for i in range(6))

    def update_single_link_response(self, waveform, lam, beta, psi, tc):
        """antenna.py:403-460: the links of ``waveform`` stored as ``detector.single_link_response``
        (complex (6, N) tensor)."""
        self.detector.single_link_response = self.detector._links(waveform, lam, beta, psi, tc)


class TDResponseModelCornish2003:
    """Time-domain single-link response (Cornish & Rubbo 2003), the reference's commented
    TDResponseModelCornish2003 (antenna.py:566-800), evaluated in a HIP kernel.

    At each time t of the TDI model's extended link grid, with the constellation at t,
    y_rs(t) = n.[H(t - L - k.x_s) - H(t - k.x_r)].n / (2 (1 - k.n_sr)) (link_rs sent from s,
    received at r; antenna.py:681-800).  The waveform container is the reference's
    ``{"t0": float, "data": (n, 2) array of [h+, hx]}`` (antenna.py:630-648), sampled at the
    data's delta_time (or the container's ``"delta_time"``); its ``t0`` is the time of the first
    sample relative to the coalescence time, so sample m sits at tc + t0 + m dt.  Between samples
    the waveform is linearly interpolated (the ``"linear"`` kernel, utils.py:83-100) and outside
    them it is zero, which is what the reference's padding in _ensure_waveform_length supplies.
    """

    def __init__(self, interpolate_kernel="linear", backend="hip"):
        if interpolate_kernel != "linear":
            raise ValueError("only the 'linear' interpolation kernel is implemented (as in the reference)")
        if backend not in BACKENDS:
            raise ValueError(f"Invalid backend: {backend}")
        self.interpolate_kernel = interpolate_kernel
        self.interpolate_kernel_length = 3
        self.backend = backend
        self.detector = None

    def init_single_link_response_model(self, detector):
        if getattr(detector.tdi_combination, "domain", None) != "td":
            raise ValueError("the time-domain response needs a time-domain TDI combination "
                             "(TDMichelsonConstantEqualArm)")
        self.detector = detector
        detector.single_link_response = None

    def _links(self, waveform_container, lam, beta, psi, tc):
        from ..core import td_links
        det = self.detector
        info = det.tdi_data.data_info
        comb = det.tdi_combination
        data = waveform_container["data"]
        if isinstance(data, torch.Tensor):
            data = data.to(torch.float64)
        else:
            data = np.asarray(data, dtype=np.float64)
        if data.ndim != 2 or data.shape[1] != 2:
            raise ValueError(f"waveform data must be (n, 2) [h+, hx], got {tuple(data.shape)}")
        wdt = float(waveform_container.get("delta_time", info.delta_time))
        orb = det.orbit_model
        t_first = info.start_time - comb.added_time_samples_number * info.delta_time
        return td_links(orb.kind, orb.arm_length, orb.rotation_initial, orb.revolution_initial, data[:, 0],
                        data[:, 1], float(tc) + float(waveform_container["t0"]), wdt, lam, beta, psi, t_first,
                        info.delta_time, comb.extended_time_series_length, **getattr(orb, "geo_params", {}))

    def get_response_jit(self, waveform_container, lam, beta, psi, tc):
        """Six real (extended_time_series_length,) GPU tensors; nothing is stored."""
        links = self._links(waveform_container, lam, beta, psi, tc)
        return tuple(links[i] for i in range(6))

    def update_single_link_response(self, waveform_container, lam, beta, psi, tc):
        """antenna.py:650-669: the links stored as ``detector.single_link_response`` (real (6, n_ext))."""
        self.detector.single_link_response = self._links(waveform_container, lam, beta, psi, tc)


class InterferometerAntenna:
    """antenna.py:117-323 with backend='hip'."""

    def __init__(self, name, tdi_data, orbit_model, response_model, tdi_combination, backend="hip", duty_cycle=None):
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        if duty_cycle is not None and not isinstance(duty_cycle, DutyCycle):
            duty_cycle = DutyCycle(*duty_cycle)
        self.duty_cycle = duty_cycle
        if isinstance(orbit_model, str):
            models = available_orbit_models(backend)
            if orbit_model not in models:
                raise ValueError(f"{orbit_model} is not a implemented orbit model. \n"
                                 f"Current available models are {[*models.keys()]}")
            orbit_model = models[orbit_model]
        elif not isinstance(orbit_model, OrbitModelBase):
            raise TypeError(f"Expected OrbitModelBase or str for orbit_model, but got {type(orbit_model).__name__}.")
        for component in (tdi_data, orbit_model, response_model, tdi_combination):
            comp_backend = getattr(component, "backend")
            if comp_backend != backend:
                raise ValueError(f"Backend mismatch: the passed class {component.__class__.__name__} uses "
                                 f"{comp_backend} as its backend. But current backend for InterferometerAntenna "
                                 f"is setting to {backend}.")
        self.name = name
        self.tdi_data = tdi_data
        self.orbit_model = orbit_model
        self.response_model = response_model
        self.tdi_combination = tdi_combination
        self.backend = backend
        self.tdi_response = None
        self._core = None
        self._core_key = None
        self.tdi_combination.init_tdi_combination_model(self)
        self.response_model.init_single_link_response_model(self)

    # -- HIP context holding this detector's grid (data enter only through the likelihood) --
    def _response_core(self, modes, reference_frequency, n_nodes, f_node_min, f_node_max):
        from ..core import LikelihoodCore
        key = (tuple(modes), reference_frequency, n_nodes, f_node_min, f_node_max)
        if self._core is None or self._core_key != key:
            core = LikelihoodCore(modes=modes, reference_frequency=reference_frequency, n_nodes=n_nodes,
                                  f_node_min=f_node_min, f_node_max=f_node_max)
            add_detector_to_core(core, self, data_channels=())
            self._core = core
            self._core_key = key
        return self._core

    def _links(self, waveform, lam, beta, psi, tc):
        """Complex (6, N) single links of a hip waveform model or a reference-style container."""
        from ..waveform.phenom import PhenomWaveformBase
        if isinstance(waveform, PhenomWaveformBase):
            row = waveform.parameter_row(lam, beta, psi, tc)
            core = self._response_core(waveform.modes, waveform.reference_frequency, waveform.n_nodes,
                                       waveform.f_node_min, waveform.f_node_max)
            return core.links(0, row[None, :])[0]
        core = self._response_core(("22",), 0.0, 2, None, None)
        return core.links_from_container(0, waveform, lam, beta, psi, tc)

    def _tdi_from_links(self, links):
        """Complex (3, N) TDI channels of this detector's combination from (6, N) links."""
        core = self._core if self._core is not None else self._response_core(("22",), 0.0, 2, None, None)
        return core.tdi_from_links(0, links)

    def update_detector_response(self, waveform, lam, beta, psi, tc):
        """antenna.py:204-229."""
        self.response_model.update_single_link_response(waveform, lam, beta, psi, tc)
        self.tdi_combination.update_tdi_response()

    def inject_signal(self, waveform, lam, beta, psi, tc):
        """antenna.py:231-271 (time-domain models add into the time-domain data)."""
        self.update_detector_response(waveform, lam, beta, psi, tc)
        chans = self.tdi_data.data_info.channels
        DataContent = namedtuple("DataContent", chans)
        resp = self.tdi_response_numpy
        if getattr(self.tdi_combination, "domain", "fd") == "td":
            self.tdi_data.add_into_td_data(DataContent(**{c: getattr(resp, c) for c in chans}))
        else:
            self.tdi_data.add_into_fd_data(DataContent(**{c: getattr(resp, c) for c in chans}))

    @property
    def tdi_response_numpy(self):
        labels = self.tdi_combination.labels
        DataContent = namedtuple("DataContent", labels)
        return DataContent(*[getattr(self.tdi_response, c).detach().cpu().numpy() for c in labels])

    @property
    def single_link_response_array(self):
        """The stored single links as a NamedTuple of complex numpy arrays (antenna.py:316-323)."""
        if self.single_link_response is None:
            raise ValueError("no single-link response: call update_detector_response first")
        DataContent = namedtuple("DataContent", LINK_NAMES)
        return DataContent(*[self.single_link_response[i].detach().cpu().numpy() for i in range(6)])


def add_detector_to_core(core, det, data_channels):
    """Register an InterferometerAntenna's grid/data/PSD with a LikelihoodCore."""
    labels = det.tdi_combination.labels
    td = det.tdi_data
    f = np.asarray(td.frequency_samples, dtype=np.float64)
    n = f.size
    data = np.zeros((3, n), np.complex128)
    psd = np.ones((3, n), np.float64)
    mask = [0, 0, 0]
    for i, c in enumerate(labels):
        if c in data_channels:
            data[i] = np.asarray(getattr(td.fd_data, c))
            psd[i] = np.asarray(getattr(td.fd_noise_power_density, c))
            mask[i] = 1
    orb = det.orbit_model
    weight = td.frequency_weights()
    uniform = td.data_info.delta_frequency is not None
    return core.add_detector(orb.kind, orb.arm_length, orb.rotation_initial, orb.revolution_initial,
                             det.tdi_combination.generation_code, det.tdi_combination.orthogonal, f, data, psd, mask,
                             df=td.data_info.delta_frequency if uniform else 0.0,
                             weight=None if uniform else weight, duty_cycle=getattr(det, "duty_cycle", None),
                             **getattr(orb, "geo_params", {}))
