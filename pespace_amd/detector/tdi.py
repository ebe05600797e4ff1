"""TDI data container and the Michelson TDI models, hip backend.

Mirrors pespace/detector/tdi.py: ``DataInformation`` (tdi.py:64-160),
``TDIChannelData`` (tdi.py:162-1376; the FD and TD parts), ``FDMichelsonConstantEqualArm``
(tdi.py:1603-1647) and ``TDMichelsonConstantEqualArm`` (tdi.py:1388-1600).  Grid, PSD and
noise are setup-time host work (numpy, as in the reference); the per-evaluation response, TDI
combination and likelihood run in the HIP library.
"""

from __future__ import annotations

import warnings
from collections import namedtuple
from dataclasses import dataclass, field

import numpy as np

from . import h5file
from .noise import FrequencyDomainNoiseModel, available_noise_models

BACKENDS = ("hip",)

# DataInformation's fields in the reference's order (tdi.py:103-118): the data_info datasets
_DATA_INFO_KEYS = ("channels", "duration", "delta_time", "start_time", "minimum_frequency", "maximum_frequency",
                   "sampling_frequency", "delta_frequency", "time_series_length", "time_samples_array",
                   "full_frequency_series_length", "full_frequency_samples_array", "frequency_mask_array",
                   "frequency_samples_array", "frequency_series_length")


def _text(v):
    """A stored string (str, bytes or a 0-d array of either) as str."""
    if isinstance(v, np.ndarray):
        v = v.item() if v.ndim == 0 else v
    return v.decode("utf-8") if isinstance(v, (bytes, np.bytes_)) else str(v)


def _texts(v):
    return tuple(_text(x) for x in np.asarray(v, dtype=object).ravel())


@dataclass(frozen=True)
class DataInformation:
    """tdi.py:64-160: rfft grid of a (duration, delta_time) series, masked to [fmin, fmax]."""

    channels: tuple
    duration: float
    delta_time: float
    start_time: float
    minimum_frequency: float
    maximum_frequency: float

    sampling_frequency: float = field(init=False)
    delta_frequency: float = field(init=False)
    time_series_length: int = field(init=False)
    full_frequency_series_length: int = field(init=False)
    full_frequency_samples_array: np.ndarray = field(init=False, repr=False)
    frequency_mask_array: np.ndarray = field(init=False, repr=False)
    frequency_samples_array: np.ndarray = field(init=False, repr=False)
    frequency_series_length: int = field(init=False)

    def __post_init__(self):
        sf = 1 / self.delta_time
        df = 1 / self.duration  # tdi.py:130
        n_t = int(self.duration // self.delta_time + 1)  # tdi.py:131
        full = np.fft.rfftfreq(n_t, self.delta_time)  # tdi.py:136
        mask = (full >= self.minimum_frequency) * (full <= self.maximum_frequency)
        f = full[mask]
        for k, v in (("sampling_frequency", sf), ("delta_frequency", df), ("time_series_length", n_t),
                     ("full_frequency_series_length", int(full.size)), ("full_frequency_samples_array", full),
                     ("frequency_mask_array", mask), ("frequency_samples_array", f),
                     ("frequency_series_length", int(f.size))):
            object.__setattr__(self, k, v)

    @property
    def time_samples_array(self):
        return np.arange(self.time_series_length) * self.delta_time + self.start_time


@dataclass(frozen=True)
class CustomGridInformation:
    """An explicit (e.g. log-spaced) frequency grid with per-bin weights df_k.

    Not in the reference, whose grids are always rfft grids; used for
    lnL = -2 sum_k df_k sum_c |d-h|^2/S_c on non-uniform grids, which reduces to
    tdi.py's df = 1/T form when df_k is constant.
    """

    channels: tuple
    frequency_samples_array: np.ndarray
    frequency_weights_array: np.ndarray

    @property
    def frequency_series_length(self):
        return int(self.frequency_samples_array.size)

    @property
    def delta_frequency(self):
        return None


class TDIChannelData:
    """FD TDI data, noise PSD and noise realizations (tdi.py:162-1376, hip backend)."""

    def __init__(self, scaling: bool = False, backend: str = "hip"):
        if backend not in BACKENDS:
            raise ValueError(f"Invalid backend: {backend}")
        self.scaling = scaling
        self.backend = backend
        self._reset()

    def _reset(self):
        self.frequency_samples = None
        self.fd_data = None
        self.time_samples = None
        self.td_data = None
        self.fd_noise_power_density = None
        self._data_info = None
        self._reset_flag = False

    @property
    def data_info(self):
        return self._data_info

    def set_data_info(self, channels, duration, delta_time, start_time, minimum_frequency, maximum_frequency):
        if self._reset_flag:
            warnings.warn("You are setting `data_info`, but you have set TDI data of current instance previously.")
        self._data_info = DataInformation(tuple(channels), duration, delta_time, start_time, minimum_frequency,
                                          maximum_frequency)

    def _store(self, data_array):
        chans = self._data_info.channels
        DataContent = namedtuple("DataContent", chans)
        self.frequency_samples = np.asarray(self._data_info.frequency_samples_array, dtype=np.float64)
        self.fd_data = DataContent(*[np.asarray(a, dtype=np.complex128) for a in data_array])
        self._reset_flag = True

    def set_fd_data_from_zero(self, channels, duration, delta_time, start_time=0.0, minimum_frequency=1e-5,
                              maximum_frequency=0.1):
        """tdi.py:585-649."""
        if self._reset_flag:
            warnings.warn("You are setting `fd_data` with zero value, but you have set TDI data previously; "
                          "the instance is reset.")
            self._reset()
        self.set_data_info(channels, duration, delta_time, start_time, minimum_frequency, maximum_frequency)
        n = self._data_info.frequency_series_length
        self._store(np.zeros((len(channels), n), dtype=np.complex128))

    def set_fd_data_from_input(self, channels, duration, delta_time, tdi_data_array, start_time=0.0,
                               minimum_frequency=1e-5, maximum_frequency=0.1):
        """tdi.py:416-519."""
        if self._reset_flag:
            warnings.warn("You are setting `fd_data` with input array, but you have set TDI data previously; "
                          "the instance is reset.")
            self._reset()
        self.set_data_info(channels, duration, delta_time, start_time, minimum_frequency, maximum_frequency)
        arr = np.asarray(tdi_data_array)
        nc, ns = arr.shape
        if nc != len(channels):
            raise ValueError(f"You set channels with {channels}, while the length of first dimension of input "
                             f"tdi_data_array is {nc}.")
        if ns != self._data_info.frequency_series_length:
            raise ValueError(f"The length of second dimension of input array is {ns}, different from "
                             f"frequency_series_length={self._data_info.frequency_series_length}.")
        self._store(arr)

    def _store_td(self, data_array):
        chans = self._data_info.channels
        DataContent = namedtuple("DataContent", chans)
        self.time_samples = np.asarray(self._data_info.time_samples_array, dtype=np.float64)
        self.td_data = DataContent(*[np.asarray(a, dtype=np.float64) for a in data_array])
        self._reset_flag = True

    def set_td_data_from_zero(self, channels, duration, delta_time, start_time=0.0, minimum_frequency=1e-5,
                              maximum_frequency=0.1):
        """tdi.py:521-583: zero time series at start_time + n delta_time, n < int(duration // delta_time + 1)."""
        if self._reset_flag:
            warnings.warn("You are setting `td_data` with zero value, but you have set TDI data previously; "
                          "the instance is reset.")
            self._reset()
        self.set_data_info(channels, duration, delta_time, start_time, minimum_frequency, maximum_frequency)
        self._store_td(np.zeros((len(channels), self._data_info.time_series_length)))

    def set_td_data_from_input(self, channels, duration, delta_time, tdi_data_array, start_time=0.0,
                               minimum_frequency=1e-5, maximum_frequency=0.1):
        """tdi.py:330-414."""
        if self._reset_flag:
            warnings.warn("You are setting `td_data` with input array, but you have set TDI data previously; "
                          "the instance is reset.")
            self._reset()
        self.set_data_info(channels, duration, delta_time, start_time, minimum_frequency, maximum_frequency)
        arr = np.asarray(tdi_data_array)
        nc, ns = arr.shape
        if nc != len(channels):
            raise ValueError(f"You set channels with {channels}, while the length of first dimension of input "
                             f"tdi_data_array is {nc}.")
        if ns != self._data_info.time_series_length:
            raise ValueError(f"The length of second dimension of input array is {ns}. It is different with the "
                             f"`time_series_length={self._data_info.time_series_length}` which is set according "
                             "to the duration and delta_time by `int(duration//delta_time + 1)`.")
        self._store_td(arr)

    def add_into_td_data(self, input):
        """Add a per-channel NamedTuple (numpy or torch) of time series into td_data (the
        reference declares this and raises NotImplementedError, tdi.py:932-940)."""
        if self.td_data is None:
            raise ValueError("no `td_data`: set it with set_td_data_from_zero / _from_input first")
        vals = self._named_arrays(input, self._data_info.time_series_length, "td_data")
        DataContent = namedtuple("DataContent", self._data_info.channels)
        self.td_data = DataContent(**{c: getattr(self.td_data, c) + vals[c] for c in self._data_info.channels})

    def _named_arrays(self, input, n, target):
        if not (isinstance(input, tuple) and hasattr(input, "_fields")):
            raise TypeError(f"Expected input to be a NamedTuple instance, but got {type(input).__name__}.")
        vals = {}
        for c in input._fields:
            a = getattr(input, c)
            if hasattr(a, "detach"):
                a = a.detach().cpu().numpy()
            vals[c] = np.asarray(a)
        if not all(v.shape == (n,) for v in vals.values()):
            raise ValueError(f"Cannot add the input NamedTuple into the `{target}`, since there is at least one "
                             "array in the input having different length with the TDI data.")
        if set(input._fields) != set(self._data_info.channels):
            raise ValueError(f"Cannot add the input NamedTuple of array into the `{target}`, since the channnels "
                             "contained by input is different with the TDI data.")
        return vals

    def set_fd_data_on_grid(self, channels, frequencies, weights, tdi_data_array=None):
        """Explicit frequency grid with per-bin weights (log-spaced grids)."""
        f = np.asarray(frequencies, dtype=np.float64)
        w = np.asarray(weights, dtype=np.float64)
        if f.shape != w.shape or f.ndim != 1:
            raise ValueError("frequencies and weights must be 1-D arrays of equal length")
        self._reset()
        self._data_info = CustomGridInformation(tuple(channels), f, w)
        arr = np.zeros((len(channels), f.size), np.complex128) if tdi_data_array is None else tdi_data_array
        self.frequency_samples = f
        DataContent = namedtuple("DataContent", channels)
        self.fd_data = DataContent(*[np.asarray(a, dtype=np.complex128) for a in arr])
        self._reset_flag = True

    def set_fd_noise_power_density_from_model(self, noise_model, **model_kwards):
        """tdi.py:771-837 (``tdi_generation`` is passed through the keyword arguments)."""
        if isinstance(noise_model, str):
            if noise_model not in available_noise_models:
                raise ValueError(f"{noise_model} is not an implemented noise model. \n"
                                 f"Current available noise models include {[*available_noise_models.keys()]}.")
            psd = available_noise_models[noise_model]
        elif isinstance(noise_model, FrequencyDomainNoiseModel):
            psd = noise_model
        else:
            raise TypeError(f"unsupported noise model {type(noise_model).__name__}")
        if self._data_info is None:
            raise ValueError("The `data_info` has not yet set.")
        if self.fd_noise_power_density is not None:
            warnings.warn("You are setting `fd_noise_power_density` for current instance, but it has been set "
                          "previously. It will be reset and updated.")
        out = psd(self._data_info.frequency_samples_array, self._data_info.channels, self.scaling, **model_kwards)
        DataContent = namedtuple("DataContent", self._data_info.channels)
        self.fd_noise_power_density = DataContent(**{c: np.asarray(out[c], np.float64)
                                                     for c in self._data_info.channels})

    def get_fd_noise_realization(self, seed=None, output_type="numpy"):
        """tdi.py:849-930: same numpy RNG call sequence (re, im per channel, DC/Nyquist zeroed)."""
        if self.fd_noise_power_density is None:
            raise ValueError("Setting `fd_noise_power_density` before generating noise realization.")
        if output_type not in ("numpy", "hip"):
            raise ValueError(f"Unknown output_type: {output_type}")
        info = self._data_info
        rng = np.random.default_rng(seed=seed)
        noise = {}
        if isinstance(info, DataInformation):
            var = 0.5 / info.delta_frequency ** 0.5
            for chan in info.channels:
                re, im = rng.normal(0, var, (2, info.full_frequency_series_length))
                re[0] = 0.0
                im[0] = 0.0
                if np.mod(info.time_series_length, 2) == 0:
                    im[-1] = 0.0
                    re[-1] = 0.0
                psd_chan = np.asarray(getattr(self.fd_noise_power_density, chan))
                noise[chan] = (re + 1j * im)[info.frequency_mask_array] * psd_chan ** 0.5
        else:  # explicit grid: per-bin variance 0.5 / sqrt(df_k)
            var = 0.5 / info.frequency_weights_array ** 0.5
            for chan in info.channels:
                re, im = rng.normal(0, 1.0, (2, info.frequency_series_length))
                psd_chan = np.asarray(getattr(self.fd_noise_power_density, chan))
                noise[chan] = (re + 1j * im) * var * psd_chan ** 0.5
        DataContent = namedtuple("DataContent", info.channels)
        return DataContent(**noise)

    def add_into_fd_data(self, input):
        """tdi.py:942-1010: add a per-channel NamedTuple (numpy or torch) into fd_data."""
        vals = self._named_arrays(input, self._data_info.frequency_series_length, "fd_data")
        DataContent = namedtuple("DataContent", self._data_info.channels)
        self.fd_data = DataContent(**{c: getattr(self.fd_data, c) + vals[c] for c in self._data_info.channels})

    # ---- files (tdi.py:1022-1292) -------------------------------------------------------
    # The reference's HDF5 layout: class_attrs/{scaling, backend, _np_fp, _np_cp},
    # data_info/<every DataInformation field>, tdi_data/<domain>/<channel>,
    # noise_data/<kind>/<channel>.  The layout logic works on a flat {path: value} tree;
    # save_to_file / recover_from_file move that tree through h5file (h5py, or the HDF5 C library
    # with h5py's datatypes).

    def _file_tree(self):
        """The datasets save_to_file writes, as {path: value} (tdi.py:1039-1088)."""
        info = self._data_info
        if not isinstance(info, DataInformation):
            raise ValueError("only rfft grids (DataInformation) have a file layout; explicit grids do not")
        tree = {"class_attrs/scaling": np.array(self.scaling), "class_attrs/backend": self.backend,
                "class_attrs/_np_fp": "float64", "class_attrs/_np_cp": "complex128"}
        for key in _DATA_INFO_KEYS:
            tree[f"data_info/{key}"] = np.array(info.channels if key == "channels" else getattr(info, key))
        if self.td_data is not None:
            for chan in info.channels:
                tree[f"tdi_data/td_data/{chan}"] = np.asarray(getattr(self.td_data, chan))
        if self.fd_data is not None:
            for chan in info.channels:
                tree[f"tdi_data/fd_data/{chan}"] = np.asarray(getattr(self.fd_data, chan))
        if self.fd_noise_power_density is not None:
            for chan in info.channels:
                tree[f"noise_data/fd_noise_power_density/{chan}"] = np.asarray(
                    getattr(self.fd_noise_power_density, chan))
        return tree

    @staticmethod
    def _from_file_tree(tree, filename="<tree>"):
        """Rebuild an instance from a dataset tree, checking every stored derived quantity of
        the grid against its recomputation (tdi.py:1120-1292).  Data of any stored backend load
        into a hip container; wavelet-domain data are not held by it and are skipped with a
        warning."""
        out = TDIChannelData(scaling=bool(np.asarray(tree["class_attrs/scaling"])), backend="hip")
        fp, cp = _text(tree["class_attrs/_np_fp"]), _text(tree["class_attrs/_np_cp"])
        assert fp == "float64" and cp == "complex128", (
            f"Failed to recover TDIChannelData from file {filename}.\n"
            "The numerical precision of stored data does not match the settings of the current environment.\n"
            f"The stored data have _np_fp={fp}, and _np_cp={cp}.\n"
            "While TDIChannelData in the current environment is initialized with _np_fp=float64, and "
            "_np_cp=complex128.")
        g = lambda k: tree[f"data_info/{k}"]
        out.set_data_info(_texts(g("channels")), float(g("duration")), float(g("delta_time")),
                          float(g("start_time")), float(g("minimum_frequency")), float(g("maximum_frequency")))
        info = out._data_info
        for key in _DATA_INFO_KEYS[6:]:
            mine, stored = getattr(info, key), np.asarray(g(key))
            same = np.array_equal(mine, stored) if np.ndim(mine) else mine == stored.item()
            assert same, (f"Failed to recover TDIChannelData from file {filename}, the `{key}` is different "
                          "from the stored value.")
        if any(k.startswith("tdi_data/wd_data/") for k in tree):
            warnings.warn(f"{filename}: `wd_data` is not held by the hip TDIChannelData and is skipped.")
        if any(k.startswith("tdi_data/td_data/") for k in tree):
            arrs = [np.asarray(tree[f"tdi_data/td_data/{c}"]) for c in info.channels]
            if any(a.shape != (info.time_series_length,) for a in arrs):
                raise ValueError(f"{filename}: stored td_data do not have time_series_length="
                                 f"{info.time_series_length} samples")
            out._store_td(arrs)
        if any(k.startswith("tdi_data/fd_data/") for k in tree):
            out._store([np.asarray(tree[f"tdi_data/fd_data/{c}"]) for c in info.channels])
        if any(k.startswith("noise_data/wd_noise_power_density/") for k in tree):
            raise NotImplementedError("wavelet-domain noise power densities are not supported")
        if any(k.startswith("noise_data/fd_noise_power_density/") for k in tree):
            DataContent = namedtuple("DataContent", info.channels)
            out.fd_noise_power_density = DataContent(**{
                c: np.asarray(tree[f"noise_data/fd_noise_power_density/{c}"], np.float64) for c in info.channels})
        out._reset_flag = True
        return out

    def save_to_file(self, filename: str) -> None:
        """tdi.py:1022-1088: the instance as an HDF5 file (fails if the file exists)."""
        h5file.write_tree(filename, self._file_tree())

    @staticmethod
    def recover_from_file(filename: str) -> "TDIChannelData":
        """tdi.py:1090-1292: an instance from a file written by save_to_file."""
        return TDIChannelData._from_file_tree(h5file.read_tree(filename), filename)

    @property
    def fd_data_numpy(self):
        return self.fd_data

    @property
    def fd_noise_power_density_numpy(self):
        return self.fd_noise_power_density

    def frequency_weights(self):
        """Per-bin df_k (constant 1/T on rfft grids)."""
        info = self._data_info
        if isinstance(info, DataInformation):
            return np.full(info.frequency_series_length, info.delta_frequency)
        return info.frequency_weights_array


class FDMichelsonConstantEqualArm:
    """FD Michelson TDI 1.5 / 2.0, A/E/T or X/Y/Z (tdi.py:1603-1647); evaluated in the HIP kernels."""

    domain = "fd"

    def __init__(self, generation="1.5", orthogonal=True, backend="hip"):
        self.generation = str(generation)
        if self.generation not in ("1.5", "2.0"):
            raise ValueError(f"Unsupported generation {self.generation}.")
        self.orthogonal = bool(orthogonal)
        self.labels = ("A", "E", "T") if self.orthogonal else ("X", "Y", "Z")
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        self.backend = backend
        self.detector = None

    @property
    def generation_code(self):
        return 15 if self.generation == "1.5" else 20

    def init_tdi_combination_model(self, detector):
        self.detector = detector
        detector.tdi_response = None

    def get_combination_jit(self, single_link_response):
        """Three channels (A, E, T or X, Y, Z) from the six links (link12, link21, link23, link32,
        link31, link13), each a complex (N,) tensor on the GPU (_tdi_jax.py:70-102)."""
        import torch
        links = torch.stack([torch.as_tensor(x) for x in single_link_response])
        out = self.detector._tdi_from_links(links)
        return tuple(out[i] for i in range(3))

    def update_tdi_response(self):
        """tdi.py:1705-1793: combine the detector's stored single links into ``tdi_response``."""
        from collections import namedtuple
        links = getattr(self.detector, "single_link_response", None)
        if links is None:
            raise ValueError("no single-link response to combine: update the single-link response first")
        out = self.detector._tdi_from_links(links)
        DataContent = namedtuple("DataContent", self.labels)
        self.detector.tdi_response = DataContent(*[out[i] for i in range(3)])


class TDMichelsonConstantEqualArm:
    """Time-domain Michelson TDI 1.5 / 2.0 for constant equal arms, A/E/T or X/Y/Z
    (tdi.py:1388-1600, marked incomplete there); evaluated in a HIP kernel.

    The single links are sampled on an extended grid that starts ``added_time_samples_number``
    samples before the data's first time sample (int(max_num_delay L // dt + 1), max_num_delay 3
    or 7, tdi.py:1436-1455); output sample i combines the links at extended sample
    i + added_time_samples_number and their delays y(t - m L), linearly interpolated.  Where the
    reference's loop (tdi.py:1482-1497) takes the base sample i + floor((max_num_delay - m) L / dt)
    with fraction 1 - frac(m L / dt), this reads at i + added - m L / dt, the delay's
    definition: the two agree except where frac(m L / dt) > frac(max_num_delay L / dt), where the
    reference's index is one sample early.
    """

    domain = "td"

    def __init__(self, generation="1.5", orthogonal=True, backend="hip"):
        self.generation = str(generation)
        if self.generation == "1.5":
            self.max_num_delay = 3
        elif self.generation == "2.0":
            self.max_num_delay = 7
        else:
            raise ValueError(f"Unsupported generation {self.generation}.")
        self.orthogonal = bool(orthogonal)
        self.labels = ("A", "E", "T") if self.orthogonal else ("X", "Y", "Z")
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend: {backend}")
        self.backend = backend
        self.added_time_samples_number = None
        self.extended_time_series_length = None
        self.detector = None

    @property
    def generation_code(self):
        return 15 if self.generation == "1.5" else 20

    def init_tdi_combination_model(self, detector):
        """tdi.py:1435-1455: sizes of the extended link grid."""
        from ..core import td_added_samples
        self.detector = detector
        detector.tdi_response = None
        info = detector.tdi_data.data_info
        if info is None or not isinstance(info, DataInformation):
            raise ValueError("the time-domain TDI needs the data's time grid: set the TDI data (set_td_data_...) "
                             "before building the detector")
        self.added_time_samples_number = td_added_samples(detector.orbit_model.arm_length, info.delta_time,
                                                          self.generation)
        self.extended_time_series_length = int(info.time_series_length + self.added_time_samples_number)

    @property
    def extended_time_samples_array(self):
        """Times of the extended link grid: start_time + (j - added) delta_time."""
        info = self.detector.tdi_data.data_info
        j = np.arange(self.extended_time_series_length, dtype=np.float64) - self.added_time_samples_number
        return info.start_time + j * info.delta_time

    def get_combination_jit(self, single_link_response):
        """Three channels (each a real (time_series_length,) GPU tensor) from the six links
        (link12, link21, link23, link32, link31, link13), each (extended_time_series_length,)."""
        import torch
        from ..core import td_tdi
        links = torch.stack([torch.as_tensor(x) for x in single_link_response])
        out = self._combine(links, td_tdi)
        return tuple(out[i] for i in range(3))

    def _combine(self, links, td_tdi):
        info = self.detector.tdi_data.data_info
        if links.shape[-1] != self.extended_time_series_length:
            raise ValueError(f"links have {links.shape[-1]} samples; the extended grid has "
                             f"{self.extended_time_series_length}")
        return td_tdi(self.detector.orbit_model.arm_length, info.delta_time, self.generation, self.orthogonal, links)

    def update_tdi_response(self):
        """tdi.py:1457-1600: combine the detector's stored single links into ``tdi_response``."""
        from ..core import td_tdi
        links = getattr(self.detector, "single_link_response", None)
        if links is None:
            raise ValueError("no single-link response to combine: update the single-link response first")
        out = self._combine(links, td_tdi)
        DataContent = namedtuple("DataContent", self.labels)
        self.detector.tdi_response = DataContent(*[out[i] for i in range(3)])
