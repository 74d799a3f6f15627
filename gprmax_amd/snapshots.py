"""Field snapshots: cell-averaged E and H over a sub-volume at one iteration
(gprMax/snapshots.py:30-130, gprMax/snapshots_ext.pyx:24-71).

A snapshot with ``time == it + 1`` is taken at the start of iteration ``it``.
Output cell (p, q, r) averages the four E edges and the two H faces of the
field cell (xs + p*dx, ys + q*dy, zs + r*dz).  The device kernel follows these
CPU semantics (the reference's GPU kernel indexes differently when xs != 0 or
the step is not 1).
"""

import struct
import sys

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

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    def write_vtk_imagedata(self, pbar=None, G=None):
        """Write the stored snapshot to ``self.filename`` as VTK ImageData
        (.vti): cell data "E-field" and "H-field", three components each,
        appended raw, each block preceded by its byte count as UInt32
        (gprMax/snapshots.py:132-167).  `pbar` (optional) is advanced by the
        bytes written; G gives the cell size."""
        if self.electric is None:
            raise ValueError('snapshot {} holds no data'.format(self.basefilename))
        e = np.ascontiguousarray(self.electric)
        h = np.ascontiguousarray(self.magnetic)
        floatname = {4: 'Float32', 8: 'Float64'}[e.dtype.itemsize]
        byteorder = 'LittleEndian' if sys.byteorder == 'little' else 'BigEndian'
        datasize = e.nbytes
        hoffset = datasize + 4
        ext = '{} {} {} {} {} {}'.format(self.xs, round_value(self.xf / self.dx), self.ys,
                                         round_value(self.yf / self.dy), self.zs, round_value(self.zf / self.dz))
        head = ('<?xml version="1.0"?>\n'
                '<VTKFile type="ImageData" version="1.0" byte_order="{}">\n'
                '<ImageData WholeExtent="{}" Origin="0 0 0" Spacing="{:.3} {:.3} {:.3}">\n'
                '<Piece Extent="{}">\n'
                '<CellData Vectors="E-field H-field">\n'
                '<DataArray type="{}" Name="E-field" NumberOfComponents="3" format="appended" offset="0" />\n'
                '<DataArray type="{}" Name="H-field" NumberOfComponents="3" format="appended" offset="{}" />\n'
                '</CellData>\n</Piece>\n</ImageData>\n<AppendedData encoding="raw">\n_').format(
            byteorder, ext, self.dx * G.dx, self.dy * G.dy, self.dz * G.dz, ext, floatname, floatname, hoffset)
        with open(self.filename, 'wb') as fh:
            fh.write(head.encode('utf-8'))
            for block in (e, h):
                fh.write(struct.pack('I', datasize))
                fh.write(block.tobytes())
                if pbar is not None:
                    pbar.update(n=4 + datasize)
            fh.write('\n</AppendedData>\n</VTKFile>'.encode('utf-8'))


def store_components(snap, Ex, Ey, Ez, Hx, Hy, Hz):
    """Store a snapshot's six [nx, ny, nz] components on `snap`: this
    package's Snapshot keeps them and builds the Paraview order; on the
    reference's Snapshot the ``electric`` and ``magnetic`` arrays are set as
    gprMax/snapshots.py:214-228 sets them."""
    if hasattr(snap, 'set_from_components'):
        snap.set_from_components(Ex, Ey, Ez, Hx, Hy, Hz)
        return
    snap.electric = np.stack([np.asarray(a, dtype=floattype) for a in (Ex, Ey, Ez)]).reshape(-1, order='F')
    snap.magnetic = np.stack([np.asarray(a, dtype=floattype) for a in (Hx, Hy, Hz)]).reshape(-1, order='F')
