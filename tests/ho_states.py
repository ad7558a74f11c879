"""Harmonic-oscillator eigenstates on a 1D grid, for closed-form checks of
the energy variance (the reference's ses::ho_eigenstate, core/src/ladder.ixx:
195-202, builds the same states from the Hermite recurrence).

psi_0 = (w/pi)^(1/4) exp(-w x^2 / 2) and
psi_{n+1} = sqrt(2w/(n+1)) x psi_n - sqrt(n/(n+1)) psi_{n-1}, carried per point
as a mantissa times exp(L) so that high levels neither underflow nor
overflow, then normalised on the grid (sum |psi|^2 dx = 1)."""
import math

import numpy as np


def ho_eigenstate(xmin, xmax, n_points, omega, level):
    h = (xmax - xmin) / n_points
    x = xmin + np.arange(n_points) * h
    log_scale = 0.25 * math.log(omega / math.pi) - 0.5 * omega * x * x
    prev = np.zeros(n_points)
    cur = np.ones(n_points)
    for n in range(level):
        nxt = math.sqrt(2.0 * omega / (n + 1)) * x * cur - math.sqrt(n / (n + 1)) * prev
        prev, cur = cur, nxt
        big = np.abs(cur) > 1e100
        if big.any():
            cur[big] *= 1e-100
            prev[big] *= 1e-100
            log_scale[big] += 100.0 * math.log(10.0)
    with np.errstate(under="ignore", over="ignore"):
        psi = cur * np.exp(log_scale)
    psi = psi / math.sqrt(float(np.sum(psi * psi)) * h)
    return psi.astype(np.complex128)
