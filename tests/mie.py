"""Mie series for a sphere (test helper).

Bohren & Huffman, "Absorption and Scattering of Light by Small Particles"
(1983), ch. 4: coefficients a_n, b_n from Riccati-Bessel functions and the
amplitude functions S1, S2.  m = None gives the perfectly conducting sphere
(a_n = psi_n'/xi_n', b_n = psi_n/xi_n).  Only magnitudes are used, so the
time convention does not matter.
"""
import numpy as np
from scipy import special


def amplitudes(x, m, theta):
    """S1, S2 at scattering angles theta (rad) for size parameter x = k a."""
    nmax = int(x + 4 * x ** (1 / 3) + 10)
    n = np.arange(1, nmax + 1)
    psi, dpsi = special.riccati_jn(nmax, x)
    chi, dchi = special.riccati_yn(nmax, x)
    xi, dxi = psi - 1j * chi, dpsi - 1j * dchi
    psi, dpsi, xi, dxi = psi[1:], dpsi[1:], xi[1:], dxi[1:]
    if m is None:
        a, b = dpsi / dxi, psi / xi
    else:
        pm, dpm = special.riccati_jn(nmax, m * x)
        pm, dpm = pm[1:], dpm[1:]
        a = (m * pm * dpsi - psi * dpm) / (m * pm * dxi - xi * dpm)
        b = (pm * dpsi - m * psi * dpm) / (pm * dxi - m * xi * dpm)
    c = (2 * n + 1) / (n * (n + 1))
    S1 = np.zeros(len(theta), complex)
    S2 = np.zeros(len(theta), complex)
    for i, t in enumerate(theta):
        mu = np.cos(t)
        pi = np.zeros(nmax + 1)
        tau = np.zeros(nmax + 1)
        pi[1] = 1.0
        for k in range(2, nmax + 1):
            pi[k] = (2 * k - 1) / (k - 1) * mu * pi[k - 1] - k / (k - 1) * pi[k - 2]
        for k in range(1, nmax + 1):
            tau[k] = k * mu * pi[k] - (k + 1) * pi[k - 1]
        S1[i] = np.sum(c * (a * pi[1:] + b * tau[1:]))
        S2[i] = np.sum(c * (a * tau[1:] + b * pi[1:]))
    return S1, S2


def rcs_eplane_db(k, radius, m, theta_s_deg):
    """Bistatic RCS (dBsm, |E_inc| = 1) for incidence along -z with E along x,
    observed at (theta_s, phi = 0): the E-plane, sigma = 4 pi |S2|^2 / k^2 at the
    scattering angle pi - theta_s."""
    S1, S2 = amplitudes(k * radius, m, np.pi - np.radians(np.asarray(theta_s_deg, float)))
    return 10 * np.log10(4 * np.pi * np.abs(S2) ** 2 / k ** 2)
