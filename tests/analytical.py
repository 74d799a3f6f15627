"""Closed-form field of a z-directed Hertzian dipole in free space.

The near + far field of a current element (doi:10.1016/0021-9991(83)90103-1),
evaluated at the staggered Yee position of each receiver component, in the
form the reference compares with (tests/analytical_solutions.py:7-158).  The
reference hard-wires a gaussiandot 1 GHz current; here the current I(t), its
time integral and its derivative are passed in, so the same solution covers
the Ricker source of the 100^3 benchmark workload:

* gaussiandot:  I = -2 z d e^{-z d^2}, int I = e^{-z d^2},
                dI/dt = 2 z (2 z d^2 - 1) e^{-z d^2}
* ricker:       I = (1 - 2 z d^2) e^{-z d^2}, int I = d e^{-z d^2},
                dI/dt = -2 z d (3 - 2 z d^2) e^{-z d^2}
with d = t - chi.
"""

import numpy as np

from gprmax_amd.constants import c, e0


def current_functions(kind, freq, amp=1.0):
    if kind in ('gaussiandot', 'gaussianprime'):
        chi, zeta = 1 / freq, 2 * np.pi ** 2 * freq ** 2

        def f(t):
            d = t - chi
            return amp * -2 * zeta * d * np.exp(-zeta * d ** 2)

        def fint(t):
            d = t - chi
            return amp * np.exp(-zeta * d ** 2)

        def fdot(t):
            d = t - chi
            return amp * 2 * zeta * (2 * zeta * d ** 2 - 1) * np.exp(-zeta * d ** 2)
    elif kind == 'ricker':
        chi, zeta = np.sqrt(2) / freq, np.pi ** 2 * freq ** 2

        def f(t):
            d = t - chi
            return amp * (1 - 2 * zeta * d ** 2) * np.exp(-zeta * d ** 2)

        def fint(t):
            d = t - chi
            return amp * d * np.exp(-zeta * d ** 2)

        def fdot(t):
            d = t - chi
            return amp * -2 * zeta * d * (3 - 2 * zeta * d ** 2) * np.exp(-zeta * d ** 2)
    else:
        raise ValueError(kind)
    return f, fint, fdot


def hertzian_dipole_fs(iterations, dt, dxdydz, rx, kind='gaussiandot', freq=1e9, amp=1.0):
    """Fields [iterations, 6] (Ex..Hz) at receiver offset ``rx`` (metres)."""
    f, fint, fdot = current_functions(kind, freq, amp)
    dx, dy, dz = dxdydz
    dl = dz
    x, y, z = rx
    t = np.arange(iterations) * dt
    out = np.zeros((iterations, 6))

    def terms(px, py, pz):
        r = np.sqrt(px ** 2 + py ** 2 + pz ** 2)
        tau = r / c
        return r, tau, fint(t - tau) * dl, f(t - tau) * dl, fdot(t - tau) * dl

    # Ex at (x + dx/2, y, z - dz/2)
    px, py, pz = x + 0.5 * dx, y, z - 0.5 * dz
    r, tau, Fi, F, Fd = terms(px, py, pz)
    out[:, 0] = (px * pz) / (4 * np.pi * e0 * r ** 5) * (3 * (Fi + tau * F) + tau ** 2 * Fd)
    # Ey at (x, y + dy/2, z - dz/2)
    px, py, pz = x, y + 0.5 * dy, z - 0.5 * dz
    r, tau, Fi, F, Fd = terms(px, py, pz)
    ratio = py / px if px != 0 else 0.0
    out[:, 1] = ratio * (px * pz) / (4 * np.pi * e0 * r ** 5) * (3 * (Fi + tau * F) + tau ** 2 * Fd)
    # Ez at (x, y, z)
    px, py, pz = x, y, z
    r, tau, Fi, F, Fd = terms(px, py, pz)
    rho2 = px ** 2 + py ** 2
    out[:, 2] = (1 / (4 * np.pi * e0 * r ** 5)) * ((2 * pz ** 2 - rho2) * (Fi + tau * F) - rho2 * tau ** 2 * Fd)
    # Hx at (x, y + dy/2, z)
    px, py, pz = x, y + 0.5 * dy, z
    r, tau, Fi, F, Fd = terms(px, py, pz)
    out[:, 3] = -(py / (4 * np.pi * r ** 3)) * (F + tau * Fd)
    # Hy at (x + dx/2, y, z)
    px, py, pz = x + 0.5 * dx, y, z
    r, tau, Fi, F, Fd = terms(px, py, pz)
    ratio = px / py if py != 0 else 0.0
    out[:, 4] = -ratio * (-(py / (4 * np.pi * r ** 3)) * (F + tau * Fd))
    return out
