"""The grid object ``G``: the data contract between model build and the solve
loop (gprMax/grid.py:80-272).

Attribute names, shapes and dtypes follow the reference so that a solve
function written against the reference's ``G`` reads this one unchanged:

* ``ID``      uint32 [6, nx+1, ny+1, nz+1], component order Ex..Hz, initialised
  to 1 (free_space);
* ``Ex``..``Hz`` float32 [nx+1, ny+1, nz+1], z fastest;
* ``updatecoeffsE/H`` float32 [nmat, 5];
* ``updatecoeffsdispersive`` complex64 [nmat, 3*maxpoles];
* ``pmls`` in the order x0, y0, z0, xmax, ymax, zmax.

The geometry helpers below only cover what the solve path's tests and
workloads need: boxes, rigidly painted voxels and the edge build of
gprMax/yee_cell_build_ext.pyx:110-257 that turns cell materials into edge IDs.
"""

from collections import OrderedDict

import numpy as np

from .constants import complextype, floattype
from .materials import Material, default_materials, process_materials
from .pml import CFS, build_pmls, default_pml_thickness
from .receivers import Rx
from .snapshots import Snapshot
from .sources import HertzianDipole, MagneticDipole, TransmissionLine, VoltageSource
from .utilities import GeneralError, courant_dt, iterations_for, round_value
from .waveforms import Waveform


class FDTDGrid:
    """Holds everything the solve loop reads and writes."""

    def __init__(self):
        self.title = ''
        self.messages = False
        self.progressbars = False
        self.nthreads = 0
        self.gpu = None
        self.nx = self.ny = self.nz = 0
        self.dx = self.dy = self.dz = 0
        self.dt = 0
        self.mode = None
        self.iterations = 0
        self.timewindow = 0
        self.pmlthickness = default_pml_thickness(10)
        self.cfs = []
        self.pmls = []
        self.pmlformulation = 'HORIPML'
        self.materials = default_materials()
        self.waveforms = []
        self.voltagesources = []
        self.hertziandipoles = []
        self.magneticdipoles = []
        self.transmissionlines = []
        self.rxs = []
        self.snapshots = []
        self.srcsteps = [0, 0, 0]
        self.rxsteps = [0, 0, 0]
        self.maxpoles = 0
        self.updatecoeffsdispersive = None
        # float32/complex64 by default; float64/complex128 reproduces the
        # reference's double-precision build (gprMax/constants.py:43-50).
        self.floattype = floattype
        self.complextype = complextype

    # ------------------------------------------------------------------ arrays
    def initialise_geometry_arrays(self):
        """solid (cell material), per-cell rigid flag and ID (gprMax/grid.py:152-167).

        The reference tracks rigidity per cell edge slot; every writer sets or
        clears all slots of a cell at once, so one flag per cell for E and one
        for H is equivalent.
        """
        self.solid = np.ones((self.nx, self.ny, self.nz), dtype=np.uint32)
        self.rigidcellE = np.zeros((self.nx, self.ny, self.nz), dtype=bool)
        self.rigidcellH = np.zeros((self.nx, self.ny, self.nz), dtype=bool)
        self.ID = np.ones((6, self.nx + 1, self.ny + 1, self.nz + 1), dtype=np.uint32)

    def initialise_field_arrays(self):
        shape = (self.nx + 1, self.ny + 1, self.nz + 1)
        for name in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'):
            setattr(self, name, np.zeros(shape, dtype=self.floattype))

    def initialise_dispersive_arrays(self):
        shape = (self.maxpoles, self.nx + 1, self.ny + 1, self.nz + 1)
        for name in ('Tx', 'Ty', 'Tz'):
            setattr(self, name, np.zeros(shape, dtype=self.complextype))

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz


# ---------------------------------------------------------------------- build
class ModelBuilder:
    """Programmatic construction of a grid, in the order run_model uses
    (gprMax/model_build_run.py:110-281).

    Usage: create, add materials / waveforms / sources / receivers / geometry,
    then call :meth:`finalise` which builds PMLs, edge IDs and coefficients.
    """

    def __init__(self, nx, ny, nz, dx, dy, dz, time_window, pml_cells=10,
                 pml_formulation='HORIPML', cfs=None, dt=None, precision='single'):
        G = FDTDGrid()
        if precision == 'double':
            G.floattype, G.complextype = np.float64, np.complex128
        elif precision != 'single':
            raise GeneralError('precision must be single or double')
        G.nx, G.ny, G.nz = int(nx), int(ny), int(nz)
        G.dx, G.dy, G.dz = float(dx), float(dy), float(dz)
        G.dt, G.mode = courant_dt(G.nx, G.ny, G.nz, G.dx, G.dy, G.dz)
        if dt is not None:
            G.dt = dt
        G.iterations, G.timewindow = iterations_for(time_window, G.dt)
        if isinstance(pml_cells, dict):
            G.pmlthickness = OrderedDict((k, int(pml_cells.get(k, 0))) for k in default_pml_thickness())
        else:
            # A thickness for every face, as without #pml_cells: a 2-D model
            # has none on its invariant axis (input_cmds_singleuse.py:145-159).
            G.pmlthickness = default_pml_thickness(int(pml_cells))
            if G.mode != '3D':
                axis = G.mode[-1]
                G.pmlthickness[axis + '0'] = G.pmlthickness[axis + 'max'] = 0
        for key, n in G.pmlthickness.items():
            size = {'x': G.nx, 'y': G.ny, 'z': G.nz}[key[0]]
            if 2 * n >= size:
                raise GeneralError('#pml_cells has too many cells for the domain size')
        G.pmlformulation = pml_formulation
        G.cfs = list(cfs) if cfs else []
        G.initialise_geometry_arrays()
        self.G = G

    # materials and waveforms ------------------------------------------------
    def add_material(self, er, se, mr, sm, name):
        m = Material(len(self.G.materials), name)
        m.er, m.se, m.mr, m.sm = float(er), float(se), float(mr), float(sm)
        if m.se == float('inf'):
            m.averagable = False
        self.G.materials.append(m)
        return m

    def add_waveform(self, kind, amp, freq, name):
        w = Waveform(ID=name, type=kind, amp=amp, freq=freq)
        self.G.waveforms.append(w)
        return w

    # sources and receivers ----------------------------------------------------
    def _place(self, src, pol, cell, waveform, start, stop):
        G = self.G
        src.polarisation = pol
        src.xcoord, src.ycoord, src.zcoord = (int(c) for c in cell)
        src.xcoordorigin, src.ycoordorigin, src.zcoordorigin = src.xcoord, src.ycoord, src.zcoord
        src.ID = '{}({},{},{})'.format(type(src).__name__, *cell)
        src.waveformID = waveform
        if start is None:
            src.start, src.stop = 0, G.timewindow
        else:
            src.start = start
            src.stop = min(stop, G.timewindow)
        src.calculate_waveform_values(G)
        return src

    def add_hertzian_dipole(self, pol, cell, waveform, start=None, stop=None):
        h = HertzianDipole()
        h.dl = {'x': self.G.dx, 'y': self.G.dy, 'z': self.G.dz}[pol]
        self.G.hertziandipoles.append(self._place(h, pol, cell, waveform, start, stop))
        return h

    def add_magnetic_dipole(self, pol, cell, waveform, start=None, stop=None):
        m = MagneticDipole()
        self.G.magneticdipoles.append(self._place(m, pol, cell, waveform, start, stop))
        return m

    def add_voltage_source(self, pol, cell, resistance, waveform, start=None, stop=None):
        v = VoltageSource()
        v.resistance = float(resistance)
        self.G.voltagesources.append(self._place(v, pol, cell, waveform, start, stop))
        return v

    def add_transmission_line(self, pol, cell, resistance, waveform, start=None, stop=None):
        """#transmission_line (gprMax/input_cmds_multiuse.py:330-383): a line of
        impedance 0 < resistance < z0 on the E edge at `cell`; its incident
        voltage and current are computed here, as the reference does when it
        reads the command."""
        from .constants import z0
        if resistance <= 0 or resistance >= z0:
            raise GeneralError('a transmission line needs a resistance greater than zero and less than '
                               'the impedance of free space ({:.2f} Ohms)'.format(float(z0)))
        t = TransmissionLine(self.G)
        t.resistance = float(resistance)
        self._place(t, pol, cell, waveform, start, stop)
        t.calculate_incident_V_I(self.G)
        self.G.transmissionlines.append(t)
        return t

    def add_rx(self, cell, name=None):
        rx = Rx(ID=name or 'Rx({},{},{})'.format(*cell), xcoord=int(cell[0]),
                ycoord=int(cell[1]), zcoord=int(cell[2]))
        rx.allocate(self.G.iterations, ftype=self.G.floattype)
        self.G.rxs.append(rx)
        return rx

    def add_snapshot(self, xs, ys, zs, xf, yf, zf, dx, dy, dz, time, filename='snap'):
        s = Snapshot(xs, ys, zs, xf, yf, zf, dx, dy, dz, time, filename)
        self.G.snapshots.append(s)
        return s

    # geometry ---------------------------------------------------------------
    def box(self, xs, xf, ys, yf, zs, zf, material, averaging=True):
        """#box over cells [xs, xf) x [ys, yf) x [zs, zf)
        (gprMax/geometry_primitives_ext.pyx:618-690)."""
        G = self.G
        averaging = averaging and material.averagable
        n = material.numID
        G.solid[xs:xf, ys:yf, zs:zf] = n
        G.rigidcellE[xs:xf, ys:yf, zs:zf] = not averaging
        G.rigidcellH[xs:xf, ys:yf, zs:zf] = not averaging
        if not averaging:
            G.ID[0, xs:xf, ys:yf + 1, zs:zf + 1] = n
            G.ID[1, xs:xf + 1, ys:yf, zs:zf + 1] = n
            G.ID[2, xs:xf + 1, ys:yf + 1, zs:zf] = n
            G.ID[3, xs:xf + 1, ys:yf, zs:zf] = n
            G.ID[4, xs:xf, ys:yf + 1, zs:zf] = n
            G.ID[5, xs:xf, ys:yf, zs:zf + 1] = n

    def cylinder(self, x1, y1, z1, x2, y2, z2, r, material, averaging=True):
        """#cylinder between face centres (x1, y1, z1) and (x2, y2, z2) in
        metres, radius r, aligned with a grid axis
        (gprMax/input_cmds_geometry.py:527-599, build_cylinder in
        gprMax/geometry_primitives_ext.pyx:692-822).

        The face centres snap to the grid; the builder then works in float32
        (its arguments are C floats) except the distance test, which is
        double: cell (i, j) is inside when
        sqrt((f32(i*dx) + 0.5*dx - x1)^2 + (f32(j*dy) + 0.5*dy - y1)^2) <= r.
        """
        G = self.G
        d = (G.dx, G.dy, G.dz)
        p1 = [round_value(float(v) / h) * h for v, h in zip((x1, y1, z1), d)]
        p2 = [round_value(float(v) / h) * h for v, h in zip((x2, y2, z2), d)]
        if r <= 0:
            raise GeneralError('the cylinder radius should be a positive value')
        f = np.float32
        p1f, p2f, df, rf = [f(v) for v in p1], [f(v) for v in p2], [f(v) for v in d], f(r)
        snap = [(round_value(float(a / h)), round_value(float(b / h))) for a, b, h in zip(p1f, p2f, df)]
        if snap[1][0] == snap[1][1] and snap[2][0] == snap[2][1]:
            axis = 0
        elif snap[0][0] == snap[0][1] and snap[2][0] == snap[2][1]:
            axis = 1
        elif snap[0][0] == snap[0][1] and snap[1][0] == snap[1][1]:
            axis = 2
        else:
            raise GeneralError('only cylinders aligned with a grid axis are supported')
        bounds = []
        for n, size in enumerate((G.nx, G.ny, G.nz)):
            a, b = sorted((p1f[n], p2f[n]))
            if n == axis:
                s, e = round_value(float(a / df[n])), round_value(float(b / df[n]))
            else:
                s, e = round_value(float((a - rf) / df[n])) - 1, round_value(float((b + rf) / df[n])) + 1
            bounds.append((max(s, 0), min(e, size)))
        mask = np.zeros((G.nx, G.ny, G.nz), dtype=bool)
        if any(e <= s for s, e in bounds):
            return
        u, v = [n for n in range(3) if n != axis]

        def centre(n):
            idx = np.arange(*bounds[n], dtype=np.float32)
            return (idx * df[n]).astype(np.float64) + 0.5 * np.float64(df[n]) - np.float64(p1f[n])

        cu, cv = centre(u), centre(v)
        inside = np.sqrt(cu[:, None] ** 2 + cv[None, :] ** 2) <= np.float64(rf)
        sl = [slice(*b) for b in bounds]
        shape = [1, 1, 1]
        shape[u], shape[v] = inside.shape
        view = mask[tuple(sl)]
        view |= inside.reshape(shape)
        averaging = averaging and material.averagable
        if averaging:
            G.solid[mask] = material.numID
            G.rigidcellE[mask] = False
            G.rigidcellH[mask] = False
        else:
            self.paint_voxels(mask, material.numID)

    def paint_voxels(self, mask, numids):
        """Rigid (non-averaged) voxels painted in (i, j, k) order with the
        semantics of build_voxel (gprMax/geometry_primitives_ext.pyx:314-376).

        ``mask`` bool [nx, ny, nz] selects cells; ``numids`` uint32 [nx, ny, nz]
        (or a scalar) gives each selected cell's material.  A voxel writes its
        four x-edges into components 0 and 3, its four y-edges into 1 and 4 and
        its four z-edges into 2 and 5.  Where voxels share an edge the one later
        in (i, j, k) order wins, which is reproduced by writing the four
        neighbour offsets in increasing voxel order.
        """
        G = self.G
        numids = np.broadcast_to(np.asarray(numids, dtype=np.uint32), mask.shape)
        G.solid[mask] = numids[mask]
        G.rigidcellE[mask] = True
        G.rigidcellH[mask] = True
        # For the x-edge (i, j, k) the touching voxels in paint order are
        # (i, j-1, k-1) < (i, j-1, k) < (i, j, k-1) < (i, j, k); a voxel at
        # (i, j', k') writes edge (i, j'+a, k'+b) for a, b in {0, 1}.  Writing
        # the offsets (a, b) = (1,1), (1,0), (0,1), (0,0) in that order makes
        # the latest voxel win.  y- and z-edges are analogous.
        order = ((1, 1), (1, 0), (0, 1), (0, 0))
        nx, ny, nz = mask.shape
        for comps, axes in (((0, 3), (1, 2)), ((1, 4), (0, 2)), ((2, 5), (0, 1))):
            for a, b in order:
                off = [0, 0, 0]
                off[axes[0]] = a
                off[axes[1]] = b
                dst = (slice(off[0], off[0] + nx), slice(off[1], off[1] + ny), slice(off[2], off[2] + nz))
                for c in comps:
                    view = G.ID[c][dst]
                    np.copyto(view, numids, where=mask)

    # finalise ---------------------------------------------------------------
    def finalise(self, edge_device=None):
        """PMLs, edge build, voltage-source materials, coefficients, receivers
        (gprMax/model_build_run.py:187-281).  With `edge_device` (a GPU
        index) the edge indices are built on that GPU (gprmax_amd.edges)."""
        G = self.G
        if any(v > 0 for v in G.pmlthickness.values()):
            if not G.cfs:
                G.cfs = [CFS()]
            build_pmls(G)
        else:
            G.pmls = []
        if edge_device is None:
            build_electric_components(G)
            build_magnetic_components(G)
        else:
            from .edges import build_components_on_device
            build_components_on_device(G, edge_device)
        pec_invariant_axis(G)
        for vs in G.voltagesources:
            vs.create_material(G)
        process_materials(G)
        for rx in G.rxs:
            if not rx.outputs:
                rx.allocate(G.iterations, ftype=G.floattype)
        return G


# ------------------------------------------------------------- edge build
def _average_material(G, names, numIDs, cache, electric):
    """Find or create the averaged material for one edge
    (gprMax/yee_cell_build_ext.pyx:31-107).

    A candidate matches when, for every constituent name, it holds that name as
    many times as the required ID does.  For H edges the reference also accepts
    any material whose ID holds both names an even number of times, which
    includes materials that hold neither; that rule is kept.
    """
    required = '+'.join(names)
    hit = cache.get(required)
    if hit is not None:
        return hit
    found = None
    for m in G.materials:
        same = all(m.ID.count(n) == required.count(n) for n in names)
        if not electric and not same:
            same = m.ID.count(names[0]) % 2 == 0 and m.ID.count(names[1]) % 2 == 0
        if same:
            found = m.numID
            break
    if found is None:
        found = len(G.materials)
        m = Material(found, required)
        m.type = 'dielectric-smoothed'
        parts = [G.materials[n] for n in numIDs]
        m.er = np.mean([p.er for p in parts], axis=0)
        m.se = np.mean([p.se for p in parts], axis=0)
        m.mr = np.mean([p.mr for p in parts], axis=0)
        m.sm = np.mean([p.sm for p in parts], axis=0)
        G.materials.append(m)
    cache[required] = found
    return found


def _edge_build(G, comp, ranges, cell_offsets, rigidcells, electric):
    """Set ID[comp] over ``ranges`` from the materials of the cells around each
    edge, in the reference's (i, j, k) loop order for new materials."""
    (i0, i1), (j0, j1), (k0, k1) = ranges
    if i1 <= i0 or j1 <= j0 or k1 <= k0:
        return
    sl = []
    rig = np.zeros((i1 - i0, j1 - j0, k1 - k0), dtype=bool)
    for (di, dj, dk) in cell_offsets:
        box = (slice(i0 + di, i1 + di), slice(j0 + dj, j1 + dj), slice(k0 + dk, k1 + dk))
        sl.append(G.solid[box])
        rig |= rigidcells[box]
    same = np.ones_like(rig)
    for s in sl[1:]:
        same &= (s == sl[0])
    target = G.ID[comp, i0:i1, j0:j1, k0:k1]
    uniform = same & ~rig
    target[uniform] = sl[0][uniform]
    mask = ~same & ~rig
    if not mask.any():
        return
    # The cell materials of each mixed edge in the reference's numID1.. order,
    # edges in C order (the reference's loop order).
    ids = np.stack([s[mask] for s in sl], axis=1).astype(np.int64)
    cache = {}
    if any('+' in m.ID for m in G.materials[:int(ids.max()) + 1]):
        # a name holding '+' makes the match depend on the constituents'
        # order: resolve edge by edge, as the reference does
        out = np.empty(len(ids), dtype=target.dtype)
        for e, row in enumerate(ids.tolist()):
            names = [G.materials[n].ID for n in row]
            out[e] = _average_material(G, names, row, cache, electric)
        target[mask] = out
        return
    # Edges holding the same materials in any order resolve to the same
    # material: the match counts each name in the '+'-joined names, which for
    # names without '+' does not depend on their order.  Each distinct set is
    # resolved once, in the order of its first edge (so materials are created
    # in the reference's order, named and averaged from that edge's order).
    key = np.sort(ids, axis=1)
    if key.max() < 1 << 16:
        packed = np.zeros(len(key), dtype=np.uint64)
        for col in range(key.shape[1]):
            packed = (packed << np.uint64(16)) | key[:, col].astype(np.uint64)
        _, first, inverse = np.unique(packed, return_index=True, return_inverse=True)
    else:
        _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    resolved = np.empty(len(first), dtype=target.dtype)
    for u in np.argsort(first, kind='stable'):
        row = ids[first[u]].tolist()
        names = [G.materials[n].ID for n in row]
        resolved[u] = _average_material(G, names, row, cache, electric)
    target[mask] = resolved[inverse.ravel()]


def pec_invariant_axis(G):
    """PEC on the two E components across the invariant axis of a 2-D model,
    at its two grid planes (gprMax/model_build_run.py:222-241)."""
    planes = {'2D TMx': (0, (1, 2)), '2D TMy': (1, (0, 2)), '2D TMz': (2, (0, 1))}
    if G.mode not in planes:
        return
    axis, comps = planes[G.mode]
    for c in comps:
        for plane in (0, 1):
            idx = [c, slice(None), slice(None), slice(None)]
            idx[1 + axis] = plane
            G.ID[tuple(idx)] = 0


def build_electric_components(G):
    """Edge IDs of Ex, Ey, Ez (gprMax/yee_cell_build_ext.pyx:110-185)."""
    nx, ny, nz = G.nx, G.ny, G.nz
    # cell offsets in the reference's numID1..numID4 order
    _edge_build(G, 0, ((0, nx), (1, ny), (1, nz)),
                ((0, 0, 0), (0, -1, 0), (0, -1, -1), (0, 0, -1)), G.rigidcellE, True)
    _edge_build(G, 1, ((1, nx), (0, ny), (1, nz)),
                ((0, 0, 0), (-1, 0, 0), (-1, 0, -1), (0, 0, -1)), G.rigidcellE, True)
    _edge_build(G, 2, ((1, nx), (1, ny), (0, nz)),
                ((0, 0, 0), (-1, 0, 0), (-1, -1, 0), (0, -1, 0)), G.rigidcellE, True)


def build_magnetic_components(G):
    """Face IDs of Hx, Hy, Hz (gprMax/yee_cell_build_ext.pyx:188-257)."""
    nx, ny, nz = G.nx, G.ny, G.nz
    _edge_build(G, 3, ((1, nx), (0, ny), (0, nz)), ((0, 0, 0), (-1, 0, 0)), G.rigidcellH, False)
    _edge_build(G, 4, ((0, nx), (1, ny), (0, nz)), ((0, 0, 0), (0, -1, 0)), G.rigidcellH, False)
    _edge_build(G, 5, ((0, nx), (0, ny), (1, nz)), ((0, 0, 0), (0, 0, -1)), G.rigidcellH, False)
