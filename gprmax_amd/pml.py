"""CFS-PML slabs and their recursive-integration coefficients
(gprMax/pml.py:30-429).

A slab is a box touching one face of the domain.  Its E and H corrections are
applied after the main Yee update, slab by slab in the order x0, y0, z0, xmax,
ymax, zmax (gprMax/grid.py:132-136).  The auxiliary Phi arrays keep the
reference's padded shapes (gprMax/pml.py:202-219) so they can be compared
element by element with a CPU solve.
"""

from collections import OrderedDict

import numpy as np

from .constants import e0, floattype, z0

BOUNDARY_IDS = ('x0', 'y0', 'z0', 'xmax', 'ymax', 'zmax')
DIRECTIONS = ('xminus', 'yminus', 'zminus', 'xplus', 'yplus', 'zplus')
DIRECTION_CODE = {name: n for n, name in enumerate(DIRECTIONS)}
FORMULATIONS = ('HORIPML', 'MRIPML')

# Polynomial orders of the named scaling profiles (gprMax/pml.py:34).
SCALING_PROFILES = {'constant': 0, 'linear': 1, 'quadratic': 2, 'cubic': 3, 'quartic': 4,
                    'quintic': 5, 'sextic': 6, 'septic': 7, 'octic': 8}


class CFSParameter:
    """One CFS parameter (alpha, kappa or sigma) with its grading profile."""

    def __init__(self, ID=None, scaling='polynomial', scalingprofile=None,
                 scalingdirection='forward', min=0, max=0):
        self.ID = ID
        self.scaling = scaling
        self.scalingprofile = scalingprofile
        self.scalingdirection = scalingdirection
        self.min = min
        self.max = max


class CFS:
    """A complex-frequency-shifted stretching term; one per PML order.

    Defaults: alpha constant 0, kappa constant 1, sigma quartic from 0 to an
    optimum maximum (gprMax/pml.py:59-69).
    """

    def __init__(self):
        self.alpha = CFSParameter(ID='alpha', scalingprofile='constant')
        self.kappa = CFSParameter(ID='kappa', scalingprofile='constant', min=1, max=1)
        self.sigma = CFSParameter(ID='sigma', scalingprofile='quartic', min=0, max=None)

    def calculate_sigmamax(self, d, er, mr):
        """Optimum sigma max, doi:10.1109/8.546249 (gprMax/pml.py:71-83)."""
        m = SCALING_PROFILES[self.sigma.scalingprofile]
        self.sigma.max = (0.8 * (m + 1)) / (z0 * d * np.sqrt(er * mr))

    @staticmethod
    def _polynomial(order, n):
        """Profile samples at whole (E) and half (H) cell positions.

        ``n`` is thickness + 1.  The n whole-cell samples are (0..n-1)/(n-1) and
        the half-cell ones are shifted by half a cell, both raised to ``order``
        (gprMax/pml.py:85-102).
        """
        samples = (np.linspace(0, (n - 1) + 0.5, num=2 * n) / (n - 1)) ** order
        return samples[0:-1:2], samples[1::2]

    def calculate_values(self, thickness, parameter, ftype=floattype):
        """E and H profile values of one parameter over the slab
        (gprMax/pml.py:104-146)."""
        Evalues = np.zeros(thickness + 1, dtype=ftype)
        Hvalues = np.zeros(thickness + 1, dtype=ftype)
        if parameter.scalingprofile == 'constant':
            Evalues += parameter.max
            Hvalues += parameter.max
        elif parameter.scaling == 'polynomial':
            Evalues, Hvalues = self._polynomial(SCALING_PROFILES[parameter.scalingprofile], thickness + 1)
            hi, lo = parameter.max, parameter.min
            Evalues = Evalues * (hi - lo) + lo
            Hvalues = Hvalues * (hi - lo) + lo
        if parameter.scalingdirection == 'reverse':
            Evalues = Evalues[::-1]
            # After reversal the half-cell samples move one place to the left.
            Hvalues = np.roll(Hvalues[::-1], -1)
        return Evalues[:-1], Hvalues[:-1]


class PML:
    """One PML slab: extent, direction of increasing absorption, coefficients.

    ``direction`` is one of xminus/yminus/zminus/xplus/yplus/zplus; the slab box
    is ``[xs, xf) x [ys, yf) x [zs, zf)`` in cells (gprMax/pml.py:149-219).
    """

    def __init__(self, G, ID=None, direction=None, xs=0, xf=0, ys=0, yf=0, zs=0, zf=0):
        self.ID = ID
        self.direction = direction
        self.xs, self.xf = xs, xf
        self.ys, self.yf = ys, yf
        self.zs, self.zf = zs, zf
        self.nx = xf - xs
        self.ny = yf - ys
        self.nz = zf - zs
        axis = direction[0]
        self.d = {'x': G.dx, 'y': G.dy, 'z': G.dz}[axis]
        self.thickness = {'x': self.nx, 'y': self.ny, 'z': self.nz}[axis]
        self.CFS = G.cfs

    @property
    def order(self):
        return len(self.CFS)

    def phi_shapes(self):
        """Shapes of (EPhi1, EPhi2, HPhi1, HPhi2) (gprMax/pml.py:202-219)."""
        o, nx, ny, nz = len(self.CFS), self.nx, self.ny, self.nz
        axis = self.direction[0]
        if axis == 'x':
            return ((o, nx + 1, ny, nz + 1), (o, nx + 1, ny + 1, nz),
                    (o, nx, ny + 1, nz), (o, nx, ny, nz + 1))
        if axis == 'y':
            return ((o, nx, ny + 1, nz + 1), (o, nx + 1, ny + 1, nz),
                    (o, nx + 1, ny, nz), (o, nx, ny, nz + 1))
        return ((o, nx, ny + 1, nz + 1), (o, nx + 1, ny, nz + 1),
                (o, nx + 1, ny, nz), (o, nx, ny + 1, nz))

    def initialise_field_arrays(self, ftype=floattype):
        s = self.phi_shapes()
        self.EPhi1 = np.zeros(s[0], dtype=ftype)
        self.EPhi2 = np.zeros(s[1], dtype=ftype)
        self.HPhi1 = np.zeros(s[2], dtype=ftype)
        self.HPhi2 = np.zeros(s[3], dtype=ftype)

    def calculate_update_coeffs(self, er, mr, G):
        """ERA..ERF and HRA..HRF, shape [order][thickness] (gprMax/pml.py:221-274).

        sigma max is computed only when it is still unset; the CFS objects are
        shared by all slabs of a grid, so the first slab built fixes it for all.
        """
        shape = (len(self.CFS), self.thickness)
        ftype = getattr(G, 'floattype', floattype)
        names = ('ERA', 'ERB', 'ERE', 'ERF', 'HRA', 'HRB', 'HRE', 'HRF')
        for name in names:
            setattr(self, name, np.zeros(shape, dtype=ftype))
        dt = G.dt
        for x, cfs in enumerate(self.CFS):
            if not cfs.sigma.max:
                cfs.calculate_sigmamax(self.d, er, mr)
            Ealpha, Halpha = cfs.calculate_values(self.thickness, cfs.alpha, ftype)
            Ekappa, Hkappa = cfs.calculate_values(self.thickness, cfs.kappa, ftype)
            Esigma, Hsigma = cfs.calculate_values(self.thickness, cfs.sigma, ftype)
            for (alpha, kappa, sigma, RA, RB, RE, RF) in (
                    (Ealpha, Ekappa, Esigma, self.ERA, self.ERB, self.ERE, self.ERF),
                    (Halpha, Hkappa, Hsigma, self.HRA, self.HRB, self.HRE, self.HRF)):
                if G.pmlformulation == 'HORIPML':
                    denom = (2 * e0 * kappa) + dt * (alpha * kappa + sigma)
                    RA[x, :] = (2 * e0 + dt * alpha) / denom
                    RB[x, :] = (2 * e0 * kappa) / denom
                    RE[x, :] = ((2 * e0 * kappa) - dt * (alpha * kappa + sigma)) / denom
                    RF[x, :] = (2 * sigma * dt) / (kappa * denom)
                elif G.pmlformulation == 'MRIPML':
                    denom = 2 * e0 + dt * alpha
                    RA[x, :] = kappa + (dt * sigma) / denom
                    RB[x, :] = (2 * e0) / denom
                    RE[x, :] = ((2 * e0) - dt * alpha) / denom
                    RF[x, :] = (2 * sigma * dt) / denom
                else:
                    raise ValueError('unknown PML formulation {!r}'.format(G.pmlformulation))


def build_pmls(G):
    """Create G.pmls in boundary order and compute their coefficients
    (gprMax/pml.py:367-429).

    The permittivity used for sigma max is the mean over the slab's first layer
    of ``G.solid`` (index ``xs``/``ys``/``zs``), as the reference takes it.
    """
    G.pmls = []
    ermap = np.array([m.er for m in sorted(G.materials, key=lambda m: m.numID)], dtype=np.float64)
    mrmap = np.array([m.mr for m in sorted(G.materials, key=lambda m: m.numID)], dtype=np.float64)
    for key, value in G.pmlthickness.items():
        if value <= 0:
            continue
        if key == 'x0':
            pml = PML(G, ID=key, direction='xminus', xf=value, yf=G.ny, zf=G.nz)
        elif key == 'xmax':
            pml = PML(G, ID=key, direction='xplus', xs=G.nx - value, xf=G.nx, yf=G.ny, zf=G.nz)
        elif key == 'y0':
            pml = PML(G, ID=key, direction='yminus', yf=value, xf=G.nx, zf=G.nz)
        elif key == 'ymax':
            pml = PML(G, ID=key, direction='yplus', ys=G.ny - value, xf=G.nx, yf=G.ny, zf=G.nz)
        elif key == 'z0':
            pml = PML(G, ID=key, direction='zminus', zf=value, xf=G.nx, yf=G.ny)
        else:
            pml = PML(G, ID=key, direction='zplus', zs=G.nz - value, xf=G.nx, yf=G.ny, zf=G.nz)
        G.pmls.append(pml)
        if key[0] == 'x':
            layer = G.solid[pml.xs, :, :]
        elif key[0] == 'y':
            layer = G.solid[:, pml.ys, :]
        else:
            layer = G.solid[:, :, pml.zs]
        # The reference sums the per-cell values one by one in C order of the
        # layer; np.cumsum accumulates the same way, so the mean rounds alike.
        averageer = _sequential_mean(ermap[layer.ravel()])
        averagemr = _sequential_mean(mrmap[layer.ravel()])
        pml.calculate_update_coeffs(averageer, averagemr, G)


def _sequential_mean(values):
    """Mean of a float64 running sum taken strictly left to right."""
    total = float(np.cumsum(values, dtype=np.float64)[-1])
    return total / values.size


def default_pml_thickness(cells=10):
    return OrderedDict((key, cells) for key in BOUNDARY_IDS)
