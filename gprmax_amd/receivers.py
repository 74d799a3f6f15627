"""Receivers: per-iteration samples of the six field components at one cell
(gprMax/receivers.py:26-88, gprMax/fields_outputs.py:40-67).

A sample is taken at the start of an iteration, before the H update, so
``outputs[c][0]`` is the initial state and ``outputs[c][n]`` the state after n
full steps.
"""

from collections import OrderedDict

import numpy as np

from .constants import COMPONENTS, floattype


class Rx:
    """Receiver output point."""

    allowableoutputs = ['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz', 'Ix', 'Iy', 'Iz']
    gpu_allowableoutputs = allowableoutputs[:-3]
    defaultoutputs = allowableoutputs[:-3]

    def __init__(self, ID=None, xcoord=None, ycoord=None, zcoord=None):
        self.ID = ID
        self.outputs = OrderedDict()
        self.xcoord, self.ycoord, self.zcoord = xcoord, ycoord, zcoord
        self.xcoordorigin, self.ycoordorigin, self.zcoordorigin = xcoord, ycoord, zcoord

    def allocate(self, iterations, components=None, ftype=floattype):
        for name in (components or self.defaultoutputs):
            self.outputs[name] = np.zeros(iterations, dtype=ftype)


def rx_coords(G):
    """int32 [nrx, 3] receiver cells in G.rxs order."""
    coords = np.zeros((len(G.rxs), 3), dtype=np.int32)
    for n, rx in enumerate(G.rxs):
        coords[n] = (rx.xcoord, rx.ycoord, rx.zcoord)
    return coords


def scatter_rx_buffer(buf, G):
    """Copy a device receiver buffer float32 [6, iterations, nrx] back into the
    receivers' output dicts (gprMax/receivers.py:71-88)."""
    for n, rx in enumerate(G.rxs):
        for c, name in enumerate(COMPONENTS):
            rx.outputs[name] = np.array(buf[c, :, n], dtype=floattype)
