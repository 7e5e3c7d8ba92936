"""Mie series for a homogeneous sphere (physics cross-check; test infrastructure).

Written from the published BHMIE method of Bohren & Huffman, "Absorption and
Scattering of Light by Small Particles" (1983), Appendix A: the logarithmic
derivative D_n(mx) by downward recurrence, the Riccati-Bessel functions psi_n
and chi_n by upward recurrence, the coefficients a_n, b_n (B&H eq. 4.88 in
logarithmic-derivative form), and the amplitude functions S1, S2 from the
angular functions pi_n, tau_n (B&H eq. 4.74).  Time convention exp(-i w t),
as in the reference solver.

The reference has no Mie code (it is not pinned by the reference): this module
is an independent physics check of the whole pipeline at the discretisation
level, with its own pins in tests (optical theorem, Rayleigh limit).
"""
from __future__ import annotations

import numpy as np


def mie_coefficients(x: float, m: complex):
    """Return (a_n, b_n) for n = 1..nstop."""
    m = complex(m)
    y = m * x
    nstop = int(np.floor(x + 4.0 * x ** (1.0 / 3.0) + 2.0))
    nmx = int(max(nstop, abs(y))) + 15
    D = np.zeros(nmx + 1, dtype=complex)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / y - 1.0 / (D[n] + n / y)
    psi0, psi1 = np.cos(x), np.sin(x)      # psi_{-1}, psi_0
    chi0, chi1 = -np.sin(x), np.cos(x)     # chi_{-1}, chi_0
    xi1 = complex(psi1, -chi1)
    a = np.zeros(nstop, dtype=complex)
    b = np.zeros(nstop, dtype=complex)
    for n in range(1, nstop + 1):
        psi = (2.0 * n - 1.0) * psi1 / x - psi0
        chi = (2.0 * n - 1.0) * chi1 / x - chi0
        xi = complex(psi, -chi)
        ta = D[n] / m + n / x
        tb = m * D[n] + n / x
        a[n - 1] = (ta * psi - psi1) / (ta * xi - xi1)
        b[n - 1] = (tb * psi - psi1) / (tb * xi - xi1)
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = complex(psi1, -chi1)
    return a, b


def amplitudes(x: float, m: complex, theta):
    """S1(theta), S2(theta) (B&H eq. 4.74)."""
    a, b = mie_coefficients(x, m)
    mu = np.cos(np.asarray(theta, dtype=float))
    S1 = np.zeros(mu.shape, dtype=complex)
    S2 = np.zeros(mu.shape, dtype=complex)
    pi0 = np.zeros_like(mu)
    pi1 = np.ones_like(mu)
    for n in range(1, len(a) + 1):
        tau = n * mu * pi1 - (n + 1) * pi0
        fn = (2.0 * n + 1.0) / (n * (n + 1.0))
        S1 += fn * (a[n - 1] * pi1 + b[n - 1] * tau)
        S2 += fn * (a[n - 1] * tau + b[n - 1] * pi1)
        pi0, pi1 = pi1, ((2.0 * n + 1.0) * mu * pi1 - (n + 1.0) * pi0) / n
    return S1, S2


def efficiencies(x: float, m: complex):
    """(Q_ext, Q_sca) from the coefficient sums (B&H eq. 4.61-4.62)."""
    a, b = mie_coefficients(x, m)
    n = np.arange(1, len(a) + 1)
    qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return qext, qsca


def mueller(x: float, m: complex, theta, k: float = None):
    """4x4 Mueller matrix of a sphere in the solver's layout and normalisation.

    The solver reports M(theta) / |k|^2 with M11 = (|S1|^2+|S2|^2)/2 and the
    Bohren-Huffman sign conventions (farfield.cu:464-508); for a sphere S3=S4=0.
    k defaults to x (unit radius).
    """
    if k is None:
        k = x
    S1, S2 = amplitudes(x, m, theta)
    a1, a2 = np.abs(S1) ** 2, np.abs(S2) ** 2
    s12 = S1 * np.conj(S2)
    M = np.zeros((4, 4, len(S1)))
    M[0, 0] = M[1, 1] = 0.5 * (a1 + a2)
    M[0, 1] = M[1, 0] = 0.5 * (a2 - a1)
    M[2, 2] = M[3, 3] = s12.real
    M[2, 3] = -s12.imag
    M[3, 2] = s12.imag
    return M / abs(k) ** 2
