"""Host-side chain statistics: integrated autocorrelation time with Sokal's
automatic windowing, autocorrelation-corrected error bars and the
independent-samples rate used for the SLMC comparison.

The reference has no estimator of its own (its tests compare plain means,
test/model_check.cpp:171-196); these are the standard definitions
(A. D. Sokal, "Monte Carlo methods in statistical mechanics", 1996, sec. 3).
"""
from __future__ import annotations

import numpy as np


def autocorrelation(x):
    """Normalised autocorrelation function rho(t), t = 0 .. n-1 (FFT)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    d = x - x.mean()
    f = np.fft.rfft(d, n=2 * n)
    acf = np.fft.irfft(f * np.conjugate(f))[:n]
    if acf[0] == 0.0:
        out = np.zeros(n)
        out[0] = 1.0
        return out
    return acf / acf[0]


def tau_int(x, c=6.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t=1}^{W} rho(t), with the
    smallest window W >= c * tau(W) (Sokal's automatic windowing)."""
    rho = autocorrelation(x)
    n = len(rho)
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return float(max(tau, 0.5))
    return float(max(tau, 0.5))


def mean_error(x, c=6.0):
    """(mean, error, tau_int) with error = sqrt(2 tau var / n)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    t = tau_int(x, c)
    var = x.var(ddof=1) if n > 1 else 0.0
    return float(x.mean()), float(np.sqrt(2.0 * t * var / n)), t


def independent_samples_per_second(x, seconds, c=6.0):
    """Measurements per second divided by 2 tau_int."""
    t = tau_int(x, c)
    return len(x) / seconds / (2.0 * t)
