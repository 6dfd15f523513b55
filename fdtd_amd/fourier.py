"""Frequency-domain post-processing on the device (API of fdtd/fourier.py).

The recorded time histories (detector traces, source voltages, loop
currents) already live in device memory; the spectra are computed there with
torch.fft in float64/complex128 and returned as device tensors. The routines
follow the reference's numpy-backend behaviour, including which sample of a
detector record it transforms (fourier.py:186-197: the first step's record
indexed [0][0][0]), the 'edge' padding and the [begin:end) frequency window
that drops the last bin (fourier.py:122-139).
"""

from math import ceil

import numpy as np
import torch

from .detectors import BlockDetector, CurrentDetector, LineDetector
from .sources import SoftArbitraryPointSource


def _as_tensor(x, device):
    if torch.is_tensor(x):
        return x.to(device=device, dtype=torch.float64)
    if isinstance(x, (list, tuple)) and x and torch.is_tensor(x[0]):
        return torch.stack([v.to(torch.float64) for v in x]).to(device)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)


def _pad_edge(x, n):
    """numpy.pad(x, (0, n), 'edge') for a 1-D tensor."""
    if n <= 0:
        return x
    return torch.cat([x, x[-1:].expand(n)])


class FrequencyRoutines:
    """FFT and impedance of recorded histories (fourier.py:18-301)."""

    def __init__(self, grid, objs):
        self.grid = grid
        self.objs = objs

    def _device(self):
        return self.grid._E.device

    def compute_padding(self, input_data, dt, freq_window_tuple=None, fft_num_bins_in_window=None,
                        fft_bin_freq_resolution=None):
        """Padding that gives the requested bin count over the window
        (fourier.py:52-120); prints the two resolutions like the reference."""
        input_length = input_data.shape[0]
        if freq_window_tuple is None:
            begin_freq = 0
            end_freq = (input_length / 2.0) / (dt * input_length)
        else:
            begin_freq, end_freq = freq_window_tuple
        end_time = input_length * dt
        if fft_num_bins_in_window is None and fft_bin_freq_resolution is not None:
            fft_num_bins_in_window = (begin_freq - end_freq) / fft_bin_freq_resolution
        elif not (fft_num_bins_in_window or fft_bin_freq_resolution):
            fft_num_bins_in_window = input_data.shape[0]
        waveform_frequency_resolution = 1.0 / end_time
        fft_bin_resolution = (1.0 / dt) / fft_num_bins_in_window
        print("Waveform data has an intrinsic resolution of: {:.2E} Hz".format(waveform_frequency_resolution))
        print("FFT bin: {:.2E} Hz".format(fft_bin_resolution))
        required_padding = ceil(fft_num_bins_in_window / ((end_freq - begin_freq) * dt)) - input_length
        return required_padding, end_time

    def compute_frequencies(self, length_with_padding, dt, freq_window_tuple=None):
        """Bin frequencies and the [begin, end) window indices (fourier.py:122-139)."""
        spectrum_freqs = torch.fft.fftfreq(length_with_padding, d=dt, dtype=torch.float64, device=self._device())
        if freq_window_tuple is None:
            return spectrum_freqs, 0, -1
        begin_freq, end_freq = freq_window_tuple
        begin_idx = int(torch.abs(spectrum_freqs - begin_freq).argmin())
        end_idx = int(torch.abs(spectrum_freqs - end_freq).argmin())
        return spectrum_freqs, begin_idx, end_idx

    def S_parameters(self, waveform, node):
        raise NotImplementedError

    def export_touchstone_s2p(self):
        raise NotImplementedError

    def FFT(self, freq_window_tuple=None, fft_num_bins_in_window=None, fft_bin_freq_resolution=None):
        """(frequencies, spectrum) of one record (fourier.py:172-214)."""
        if self.grid.time_steps_passed == 0:
            return [], []
        o = self.objs
        if torch.is_tensor(o) or isinstance(o, np.ndarray):
            input_data = o
        elif isinstance(o, (LineDetector, BlockDetector)):
            input_data = o.E[:][0][0][0]
        elif isinstance(o, SoftArbitraryPointSource):
            input_data = o.source_voltage[:][0][0][0]
        elif isinstance(o, CurrentDetector):
            input_data = o.I[:][0][0][0]
        else:
            raise ValueError("Sorry, FFT can't yet interpret the argument given.")
        input_data = _as_tensor(input_data, self._device()).reshape(-1)
        required_padding, _ = self.compute_padding(input_data, self.grid.time_step,
                                                   freq_window_tuple=freq_window_tuple,
                                                   fft_num_bins_in_window=fft_num_bins_in_window,
                                                   fft_bin_freq_resolution=fft_bin_freq_resolution)
        input_data = _pad_edge(input_data, required_padding)
        spectrum = torch.fft.fft(input_data)
        freqs, b, e = self.compute_frequencies(input_data.shape[0], self.grid.time_step,
                                               freq_window_tuple=freq_window_tuple)
        return freqs[b:e], spectrum[b:e]

    def impedance(self, freq_window_tuple=None, fft_num_bins_in_window=None, fft_bin_freq_resolution=None):
        """(frequencies, V/I spectrum) of a soft source, or of a (detector,
        current detector) pair (fourier.py:216-266); 0 where the current
        spectrum is 0."""
        if self.grid.time_steps_passed == 0:
            return [], []
        o = self.objs
        dev = self._device()
        if isinstance(o, SoftArbitraryPointSource):
            voltage = o.trace("source_voltage").to(torch.float64)
            current = o.current_detector.trace("I").to(torch.float64)
        elif isinstance(o, tuple) and isinstance(o[0], (BlockDetector, LineDetector)) \
                and isinstance(o[1], CurrentDetector):
            voltage = _as_tensor(o[0].E, dev)[:][0, 0, 0]
            current = o[1].trace("I").to(torch.float64)
        else:
            raise ValueError("Sorry, FFT can't yet interpret the argument given.")
        required_padding, _ = self.compute_padding(voltage, self.grid.time_step,
                                                   freq_window_tuple=freq_window_tuple,
                                                   fft_num_bins_in_window=fft_num_bins_in_window,
                                                   fft_bin_freq_resolution=fft_bin_freq_resolution)
        voltage = _pad_edge(voltage.reshape(self.grid.time_steps_passed), required_padding)
        current = _pad_edge(current.reshape(self.grid.time_steps_passed), required_padding)
        vs = torch.fft.fft(voltage)
        cs = torch.fft.fft(current)
        z = torch.where(cs != 0, vs / torch.where(cs != 0, cs, torch.ones_like(cs)), torch.zeros_like(vs))
        freqs, b, e = self.compute_frequencies(voltage.shape[0], self.grid.time_step,
                                               freq_window_tuple=freq_window_tuple)
        return freqs[b:e], z[b:e]
