"""Sources of the solve path: Hertzian dipoles, magnetic dipoles, voltage
sources (gprMax/sources.py:33-283) and transmission lines
(gprMax/sources.py:286-451).

Waveforms are sampled on the host once per iteration, as the reference does
(gprMax/sources.py:47-68).  For the device each source becomes one row of
``srcinfo1 = (i, j, k, pol)`` (int32), two float32 factors and a float32 wave
table, plus the iteration window in which it is active.
"""

from copy import deepcopy

import numpy as np

from .constants import IDlookup, c, floattype
from .utilities import round_value

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
        return active_window(self, G)


def active_window(src, G):
    """First and last iteration at which the CPU path applies source `src`
    (any object with ``start`` and ``stop``, the reference's included).

    The CPU update runs only when ``start <= it*dt <= stop``
    (gprMax/sources.py:92,174,213); the window is evaluated with the same
    float64 products so the device gating matches it exactly.
    """
    active = [it for it in range(G.iterations) if src.start <= it * G.dt <= src.stop]
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


class TransmissionLine(Source):
    """A one-dimensional line of impedance ``resistance`` attached to an E
    edge (gprMax/sources.py:286-451): a one-way injector at cell ``srcpos``
    drives the line, whose cell ``antpos`` exchanges voltage and current with
    the grid; an absorbing condition terminates the other end.

    The line arithmetic keeps the reference's operand types: its cell length
    ``dl`` comes from ``np.sqrt`` (a NumPy float64), so under NumPy 2
    promotion every line coefficient is a float64 scalar applied to float32
    differences, rounded to float32 when stored.
    """

    def __init__(self, G):
        super().__init__()
        self.resistance = None
        self.abcv0 = 0
        self.abcv1 = 0
        self.dl = np.sqrt(3) * c * G.dt
        # a long line first, to record the incident voltage and current
        self.nl = round_value(0.667 * G.iterations)
        self.srcpos = 5
        self.antpos = 10
        ftype = getattr(G, 'floattype', floattype)
        self.voltage = np.zeros(self.nl, dtype=ftype)
        self.current = np.zeros(self.nl, dtype=ftype)
        self.Vtotal = np.zeros(G.iterations, dtype=ftype)
        self.Itotal = np.zeros(G.iterations, dtype=ftype)

    def coefficients(self, G):
        """(R c dt/dl, (1/R) c dt/dl, c dt/dl, ABC factor), float64."""
        cdt_dl = c * G.dt / self.dl
        return (self.resistance * cdt_dl, (1 / self.resistance) * cdt_dl, cdt_dl,
                (c * G.dt - self.dl) / (c * G.dt + self.dl))

    def calculate_incident_V_I(self, G):
        """Run the long line alone for every iteration, recording the voltage
        and current at the antenna cell, then keep only the cells up to it
        (gprMax/sources.py:335-358).  The kept cells carry the long run's
        final state into the solve, as in the reference."""
        ftype = self.voltage.dtype
        self.Vinc = np.zeros(G.iterations, dtype=ftype)
        self.Iinc = np.zeros(G.iterations, dtype=ftype)
        for it in range(G.iterations):
            self.Iinc[it] = self.current[self.antpos]
            self.Vinc[it] = self.voltage[self.antpos]
            self.update_current(it, G)
            self.update_voltage(it, G)
        self.nl = self.antpos + 1

    def update_abc(self, G):
        """First-order absorbing end at cell 0 (gprMax/sources.py:360-372)."""
        h = self.coefficients(G)[3]
        self.voltage[0] = h * (self.voltage[1] - self.abcv0) + self.abcv1
        self.abcv0 = self.voltage[0]
        self.abcv1 = self.voltage[1]

    def update_voltage(self, it, G):
        """Voltages from the current differences, the injector, the absorbing
        end (gprMax/sources.py:374-388)."""
        cV, _, cW, _ = self.coefficients(G)
        n = self.nl
        self.voltage[1:n] -= cV * (self.current[1:n] - self.current[:n - 1])
        self.voltage[self.srcpos] += cW * self.waveformvalues_wholestep[it]
        self.update_abc(G)

    def update_current(self, it, G):
        """Currents from the voltage differences and the injector
        (gprMax/sources.py:390-402)."""
        _, cI, _, _ = self.coefficients(G)
        n = self.nl
        self.current[:n - 1] -= cI * (self.voltage[1:n] - self.voltage[:n - 1])
        self.current[self.srcpos - 1] += cI * self.waveformvalues_halfstep[it]


def device_tl_args(tl, G):
    """Arguments of gpx_add_transmission_line for line ``tl`` (this module's
    TransmissionLine or the reference's: only their attributes are read):
    info int32[9], coefficients float64[4], cell sizes float32[3], the whole-
    and half-step waveforms, and the line state (voltage, current, ABC
    memory)."""
    pol = POLARISATION[tl.polarisation]
    first, last = active_window(tl, G)
    info = np.array([tl.xcoord, tl.ycoord, tl.zcoord, pol, first, last, tl.nl, tl.srcpos, tl.antpos],
                    dtype=np.int32)
    cdt_dl = c * G.dt / tl.dl            # as TransmissionLine.coefficients
    coef = np.array([tl.resistance * cdt_dl, (1 / tl.resistance) * cdt_dl, cdt_dl,
                     (c * G.dt - tl.dl) / (c * G.dt + tl.dl)], dtype=np.float64)
    d = (G.dx, G.dy, G.dz)
    loop = [d[a] for a in range(3) if a != pol]
    sizes = np.array([d[pol]] + loop, dtype=np.float32)
    n = tl.nl
    abc = np.array([tl.abcv0, tl.abcv1], dtype=np.float32)
    return (info, coef, sizes, np.ascontiguousarray(tl.waveformvalues_wholestep, dtype=np.float32),
            np.ascontiguousarray(tl.waveformvalues_halfstep, dtype=np.float32),
            np.ascontiguousarray(tl.voltage[:n], dtype=np.float32),
            np.ascontiguousarray(tl.current[:n], dtype=np.float32), abc)


SOURCE_KINDS = ('hertzian', 'magnetic', 'voltage')


def device_source_table(sources, G, kind=None):
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

    ``kind`` ('hertzian', 'magnetic' or 'voltage') names the list the sources
    come from, so the reference's own source objects pack too; without it the
    kind follows this module's classes.
    """
    n = len(sources)
    info1 = np.zeros((n, 6), dtype=np.int32)
    factors = np.zeros((n, 2), dtype=np.float32)
    waves = np.zeros((n, G.iterations), dtype=np.float32)
    inv_vol = 1 / (G.dx * G.dy * G.dz)
    steps = {'x': G.dx, 'y': G.dy, 'z': G.dz}
    for s, src in enumerate(sources):
        pol = POLARISATION[src.polarisation]
        first, last = active_window(src, G)
        info1[s] = (src.xcoord, src.ycoord, src.zcoord, pol, first, last)
        k = kind or ('hertzian' if isinstance(src, HertzianDipole) else
                     'magnetic' if isinstance(src, MagneticDipole) else
                     'voltage' if isinstance(src, VoltageSource) else None)
        if k == 'hertzian':
            factors[s] = (src.dl, inv_vol)
            waves[s] = src.waveformvalues_wholestep
        elif k == 'magnetic':
            factors[s] = (inv_vol, 0)
            waves[s] = src.waveformvalues_halfstep
        elif k == 'voltage':
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
