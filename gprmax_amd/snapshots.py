"""Field snapshots: cell-averaged E and H over a sub-volume at one iteration
(gprMax/snapshots.py:30-130, gprMax/snapshots_ext.pyx:24-71).

A snapshot with ``time == it + 1`` is taken at the start of iteration ``it``.
Output cell (p, q, r) averages the four E edges and the two H faces of the
field cell (xs + p*dx, ys + q*dy, zs + r*dz).  The device kernel follows these
CPU semantics (the reference's GPU kernel indexes differently when xs != 0 or
the step is not 1).
"""

import numpy as np

from .constants import floattype
from .utilities import round_value


class Snapshot:
    """One requested snapshot volume and its iteration."""

    def __init__(self, xs=None, ys=None, zs=None, xf=None, yf=None, zf=None,
                 dx=None, dy=None, dz=None, time=None, filename=None):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.xf, self.yf, self.zf = xf, yf, zf
        self.dx, self.dy, self.dz = dx, dy, dz
        self.nx = round_value((self.xf - self.xs) / self.dx)
        self.ny = round_value((self.yf - self.ys) / self.dy)
        self.nz = round_value((self.zf - self.zs) / self.dz)
        self.time = time
        self.basefilename = filename
        self.electric = None
        self.magnetic = None

    def set_from_components(self, Ex, Ey, Ez, Hx, Hy, Hz):
        """Store six [nx, ny, nz] component arrays in the Paraview order
        (gprMax/snapshots.py:128-130)."""
        self.fields = tuple(np.asarray(a, dtype=floattype) for a in (Ex, Ey, Ez, Hx, Hy, Hz))
        self.electric = np.stack(self.fields[:3]).reshape(-1, order='F')
        self.magnetic = np.stack(self.fields[3:]).reshape(-1, order='F')
