"""Scalar pulse shapes (fdtd/waveforms.py). Evaluated on the host with Python
``math`` so that per-step source values carry the reference's exact bits."""

from math import cos, exp, log, sin, sqrt

import numpy as np


def hanning(f, t, n):
    """Hanning-windowed sine with frequency f in 1/steps and n cycles
    (fdtd/waveforms.py:8-9; no 2*pi factor, as in the reference)."""
    return (1 / 2) * (1 - cos(f * t / n)) * (sin(f * t))


fwhm_constant = 2.0 * sqrt(2.0 * log(2))


def normalized_gaussian_pulse(x, fwhm, center=0.0):
    """Gaussian with peak 1 at `center` and the given full width at half
    maximum (fdtd/waveforms.py:34-46)."""
    sigma = fwhm / fwhm_constant
    return np.exp(-(((x - center) ** 2.0) / (2.0 * (sigma ** 2.0))))


def normalized_gaussian_derivative_pulse(x, fwhm, center=0.0):
    """Derivative-of-Gaussian pulse normalised to peak 1 (fdtd/waveforms.py:49-53)."""
    sigma = fwhm / fwhm_constant
    return (exp((1.0 / 2.0) - ((x - center) ** 2.0) / (2.0 * sigma ** 2.0)) * (x - center)) / sigma
