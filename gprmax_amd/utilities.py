"""Host helpers the solve path depends on.

Only the numeric rules that fix model inputs are kept here: integer rounding of
coordinates, the floor of the time step, and the Courant-limited time step
itself.
"""

import decimal

import numpy as np

from .constants import c


class GeneralError(Exception):
    """Raised for invalid models, invalid solver states and HIP failures."""


def round_value(value, decimalplaces=0):
    """Round as the reference does (gprMax/utilities.py:116-136).

    With ``decimalplaces == 0`` the value is rounded to the nearest integer with
    halves going down.  Otherwise it is floored to ``decimalplaces`` decimal
    places, which is how the time step is kept under the Courant limit.
    """
    exact = decimal.Decimal(value)
    if decimalplaces == 0:
        return int(exact.quantize(decimal.Decimal(1), rounding=decimal.ROUND_HALF_DOWN))
    step = decimal.Decimal(1).scaleb(-decimalplaces)
    return float(exact.quantize(step, rounding=decimal.ROUND_FLOOR))


def courant_dt(nx, ny, nz, dx, dy, dz):
    """Time step at the Courant limit, floored as in
    gprMax/input_cmds_singleuse.py:144-166.

    Returns ``(dt, mode)`` where mode is '3D', '2D TMx', '2D TMy' or '2D TMz'.
    """
    if nx == 1:
        raw, mode = 1 / (c * np.sqrt((1 / dy) * (1 / dy) + (1 / dz) * (1 / dz))), '2D TMx'
    elif ny == 1:
        raw, mode = 1 / (c * np.sqrt((1 / dx) * (1 / dx) + (1 / dz) * (1 / dz))), '2D TMy'
    elif nz == 1:
        raw, mode = 1 / (c * np.sqrt((1 / dx) * (1 / dx) + (1 / dy) * (1 / dy))), '2D TMz'
    else:
        raw = 1 / (c * np.sqrt((1 / dx) * (1 / dx) + (1 / dy) * (1 / dy) + (1 / dz) * (1 / dz)))
        mode = '3D'
    # One decimal place fewer than the default decimal context precision (28).
    return round_value(float(raw), decimalplaces=decimal.getcontext().prec - 1), mode


def iterations_for(time_window, dt):
    """Iteration count for a time window (gprMax/input_cmds_singleuse.py:184-207).

    An int is taken as the iteration count itself; a float in seconds gives
    ``ceil(tw / dt) + 1`` iterations.  Returns ``(iterations, timewindow)``.
    """
    if isinstance(time_window, (int, np.integer)):
        return int(time_window), (int(time_window) - 1) * dt
    tw = float(time_window)
    if tw <= 0:
        raise GeneralError('time window must be greater than zero')
    return int(np.ceil(tw / dt)) + 1, tw
