"""Sources of the solve path: Hertzian dipoles, magnetic dipoles and voltage
sources (gprMax/sources.py:33-283).

Waveforms are sampled on the host once per iteration, as the reference does
(gprMax/sources.py:47-68).  For the device each source becomes one row of
``srcinfo1 = (i, j, k, pol)`` (int32), two float32 factors and a float32 wave
table, plus the iteration window in which it is active.
"""

from copy import deepcopy

import numpy as np

from .constants import IDlookup, floattype

POLARISATION = {'x': 0, 'y': 1, 'z': 2}


class Source:
    """A point source at a Yee edge (electric) or face (magnetic)."""

    def __init__(self):
        self.ID = None
        self.polarisation = None
        self.xcoord = self.ycoord = self.zcoord = None
        self.xcoordorigin = self.ycoordorigin = self.zcoordorigin = None
        self.start = None
        self.stop = None
        self.waveformID = None

    def calculate_waveform_values(self, G):
        """Sample the waveform on whole and half steps inside [start, stop]
        (gprMax/sources.py:47-68)."""
        ftype = getattr(G, 'floattype', floattype)
        self.waveformvalues_wholestep = np.zeros(G.iterations, dtype=ftype)
        self.waveformvalues_halfstep = np.zeros(G.iterations, dtype=ftype)
        waveform = next(w for w in G.waveforms if w.ID == self.waveformID)
        for it in range(G.iterations):
            time = G.dt * it
            if self.start <= time <= self.stop:
                time -= self.start
                self.waveformvalues_wholestep[it] = waveform.calculate_value(time, G.dt)
                self.waveformvalues_halfstep[it] = waveform.calculate_value(time + 0.5 * G.dt, G.dt)

    def active_window(self, G):
        """First and last iteration at which the CPU path applies the source.

        The CPU update runs only when ``start <= it*dt <= stop``
        (gprMax/sources.py:92,174,213); the window is evaluated with the same
        float64 products so the device gating matches it exactly.
        """
        active = [it for it in range(G.iterations) if self.start <= it * G.dt <= self.stop]
        if not active:
            return 1, 0
        return active[0], active[-1]


class HertzianDipole(Source):
    """Additive electric current element of length dl (gprMax/sources.py:156-193)."""

    def __init__(self):
        super().__init__()
        self.dl = None


class MagneticDipole(Source):
    """Additive magnetic current element (gprMax/sources.py:196-232)."""


class VoltageSource(Source):
    """Resistive (R > 0) or hard (R == 0) voltage source
    (gprMax/sources.py:71-153)."""

    def __init__(self):
        super().__init__()
        self.resistance = None

    def create_material(self, G):
        """Add the source's conductance to a copy of the underlying material at
        its edge (gprMax/sources.py:122-153)."""
        if self.resistance == 0:
            return
        i, j, k = self.xcoord, self.ycoord, self.zcoord
        comp = IDlookup['E' + self.polarisation]
        base = next(m for m in G.materials if m.numID == G.ID[comp, i, j, k])
        new = deepcopy(base)
        new.ID = base.ID + '+' + self.ID
        new.numID = len(G.materials)
        new.averagable = False
        new.type += ',\nvoltage-source'
        if self.polarisation == 'x':
            new.se += G.dx / (self.resistance * G.dy * G.dz)
        elif self.polarisation == 'y':
            new.se += G.dy / (self.resistance * G.dx * G.dz)
        else:
            new.se += G.dz / (self.resistance * G.dx * G.dy)
        G.ID[comp, i, j, k] = new.numID
        G.materials.append(new)


def device_source_table(sources, G):
    """Pack one kind of sources for the device.

    Returns ``(info1 int32[n,6], factors float32[n,2], waves float32[n,iterations])``.
    ``info1`` rows are ``(i, j, k, pol, it_first, it_last)``.  ``factors`` hold
    the float32 constants the CPU path multiplies by, in the order it applies
    them (gprMax/sources.py:94-120, 176-193, 215-232):

    * Hertzian dipole: ``(dl, 1/(dx*dy*dz))`` on the whole-step wave;
    * magnetic dipole: ``(1/(dx*dy*dz), 0)`` on the half-step wave;
    * resistive voltage source: ``(1/(R*d_a*d_b), 0)`` on the whole-step wave;
    * hard voltage source: ``(d_pol, 0)`` on the half-step wave, ``R == 0``
      marked by ``pol + 4`` in column 3.
    """
    n = len(sources)
    info1 = np.zeros((n, 6), dtype=np.int32)
    factors = np.zeros((n, 2), dtype=np.float32)
    waves = np.zeros((n, G.iterations), dtype=np.float32)
    inv_vol = 1 / (G.dx * G.dy * G.dz)
    steps = {'x': G.dx, 'y': G.dy, 'z': G.dz}
    for s, src in enumerate(sources):
        pol = POLARISATION[src.polarisation]
        first, last = src.active_window(G)
        info1[s] = (src.xcoord, src.ycoord, src.zcoord, pol, first, last)
        if isinstance(src, HertzianDipole):
            factors[s] = (src.dl, inv_vol)
            waves[s] = src.waveformvalues_wholestep
        elif isinstance(src, MagneticDipole):
            factors[s] = (inv_vol, 0)
            waves[s] = src.waveformvalues_halfstep
        elif isinstance(src, VoltageSource):
            if src.resistance != 0:
                other = [steps[a] for a in 'xyz' if a != src.polarisation]
                factors[s] = (1 / (src.resistance * other[0] * other[1]), 0)
                waves[s] = src.waveformvalues_wholestep
            else:
                info1[s, 3] = pol + 4
                factors[s] = (steps[src.polarisation], 0)
                waves[s] = src.waveformvalues_halfstep
        else:
            raise TypeError('unsupported source type {}'.format(type(src).__name__))
    return info1, factors, waves
