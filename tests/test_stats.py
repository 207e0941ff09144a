"""Host statistics (universal_classical_montecarlo_amd/stats.py) against closed
forms: for an AR(1) process x_t = rho x_{t-1} + noise the integrated
autocorrelation time is (1 + rho) / (2 (1 - rho))."""
import numpy as np
import pytest

from universal_classical_montecarlo_amd import stats


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9])
def test_tau_int_ar1(rho):
    rng = np.random.default_rng(1)
    n = 400000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0]
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    tau = stats.tau_int(x)
    exact = (1 + rho) / (2 * (1 - rho))
    assert tau == pytest.approx(exact, rel=0.06)


def test_mean_error_covers_truth():
    rng = np.random.default_rng(2)
    x = rng.normal(loc=3.0, size=20000)
    m, err, tau = stats.mean_error(x)
    assert abs(m - 3.0) < 5 * err
    assert err == pytest.approx(1 / np.sqrt(20000), rel=0.1)
