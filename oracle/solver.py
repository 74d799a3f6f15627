"""CPU solve loop of the oracle -- TEST INFRASTRUCTURE.

Restates gprMax/model_build_run.py:408-474 (solve_cpu): per iteration, store
receivers and transmission-line outputs, store snapshots, H update, H-PML per
slab, transmission lines then magnetic sources, E update (or dispersive part
A), E-PML per slab, voltage sources, transmission lines, Hertzian dipoles,
dispersive part B.  Point sources use NumPy scalar arithmetic exactly as
gprMax/sources.py:81-232 writes it, so their rounding is the reference's.
Transmission lines (gprMax/sources.py:286-451) are restated element by
element with explicit casts: float64 products of float32 differences,
rounded to float32 when stored (the NumPy 2 promotion of the reference's
expressions, whose line coefficients are NumPy float64 scalars).  Against the
reference's outputs these line functions are parity unpinned (no output file
of the reference holds a line); they are pinned against the host's
vectorised incident run and the matched-line relation V = R I.
"""

import os
import time

import numpy as np

from gprmax_amd.constants import IDlookup, c
from gprmax_amd.pml import DIRECTION_CODE

from . import load, ptr


def default_threads():
    return int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))


def _pml_call(lib, G, pml, field, nthreads):
    fn = lib.funcs['pml_update']
    coeffs = G.updatecoeffsE if field == 0 else G.updatecoeffsH
    if field == 0:
        phi1, phi2, R = pml.EPhi1, pml.EPhi2, (pml.ERA, pml.ERB, pml.ERE, pml.ERF)
    else:
        phi1, phi2, R = pml.HPhi1, pml.HPhi2, (pml.HRA, pml.HRB, pml.HRE, pml.HRF)
    form = 0 if G.pmlformulation == 'HORIPML' else 1
    fn(field, form, len(pml.CFS), DIRECTION_CODE[pml.direction],
       pml.xs, pml.xf, pml.ys, pml.yf, pml.zs, pml.zf, nthreads,
       G.nx, G.ny, G.nz, ptr(coeffs), ptr(G.ID),
       ptr(G.Ex), ptr(G.Ey), ptr(G.Ez), ptr(G.Hx), ptr(G.Hy), ptr(G.Hz),
       ptr(phi1), ptr(phi2), *(ptr(r) for r in R), lib.real(pml.d))


def hertzian_dipole_update(src, it, G):
    """gprMax/sources.py:163-193."""
    if it * G.dt >= src.start and it * G.dt <= src.stop:
        i, j, k = src.xcoord, src.ycoord, src.zcoord
        c = 'E' + src.polarisation
        F = getattr(G, c)
        F[i, j, k] -= (G.updatecoeffsE[G.ID[IDlookup[c], i, j, k], 4]
                       * src.waveformvalues_wholestep[it] * src.dl * (1 / (G.dx * G.dy * G.dz)))


def magnetic_dipole_update(src, it, G):
    """gprMax/sources.py:202-232."""
    if it * G.dt >= src.start and it * G.dt <= src.stop:
        i, j, k = src.xcoord, src.ycoord, src.zcoord
        c = 'H' + src.polarisation
        F = getattr(G, c)
        F[i, j, k] -= (G.updatecoeffsH[G.ID[IDlookup[c], i, j, k], 4]
                       * src.waveformvalues_halfstep[it] * (1 / (G.dx * G.dy * G.dz)))


def voltage_source_update(src, it, G):
    """gprMax/sources.py:81-120."""
    if it * G.dt >= src.start and it * G.dt <= src.stop:
        i, j, k = src.xcoord, src.ycoord, src.zcoord
        c = 'E' + src.polarisation
        F = getattr(G, c)
        d = {'x': (G.dx, G.dy, G.dz), 'y': (G.dy, G.dx, G.dz), 'z': (G.dz, G.dx, G.dy)}[src.polarisation]
        if src.resistance != 0:
            F[i, j, k] -= (G.updatecoeffsE[G.ID[IDlookup[c], i, j, k], 4]
                           * src.waveformvalues_wholestep[it] * (1 / (src.resistance * d[1] * d[2])))
        else:
            F[i, j, k] = -src.waveformvalues_halfstep[it] / d[0]


def _tl_coefficients(tl, G):
    cdt_dl = float(c) * float(G.dt) / float(tl.dl)
    h = (float(c) * float(G.dt) - float(tl.dl)) / (float(c) * float(G.dt) + float(tl.dl))
    return float(tl.resistance) * cdt_dl, (1.0 / float(tl.resistance)) * cdt_dl, cdt_dl, h


def tl_update_current(tl, it, G):
    """gprMax/sources.py:390-402, one line cell at a time."""
    _, cI, _, _ = _tl_coefficients(tl, G)
    V, I = tl.voltage, tl.current
    for m in range(tl.nl - 1):
        d = np.float32(V[m + 1]) - np.float32(V[m])
        I[m] = np.float32(float(I[m]) - cI * float(d))
    q = tl.srcpos - 1
    I[q] = np.float32(float(I[q]) + cI * float(tl.waveformvalues_halfstep[it]))


def tl_update_voltage(tl, it, G):
    """gprMax/sources.py:360-388, one line cell at a time, then the absorbing end."""
    cV, _, cW, h = _tl_coefficients(tl, G)
    V, I = tl.voltage, tl.current
    for m in range(1, tl.nl):
        d = np.float32(I[m]) - np.float32(I[m - 1])
        V[m] = np.float32(float(V[m]) - cV * float(d))
    q = tl.srcpos
    V[q] = np.float32(float(V[q]) + cW * float(tl.waveformvalues_wholestep[it]))
    a0, a1 = np.float32(tl.abcv0), np.float32(tl.abcv1)
    V[0] = np.float32(h * float(np.float32(V[1]) - a0) + float(a1))
    tl.abcv0 = V[0]
    tl.abcv1 = V[1]


def tl_loop_current(tl, G):
    """Current of the H loop around the line's edge (gprMax/grid.py:413-461)."""
    x, y, z = tl.xcoord, tl.ycoord, tl.zcoord
    f = np.float32
    Hx, Hy, Hz = G.Hx, G.Hy, G.Hz
    if tl.polarisation == 'x':
        if y == 0 or z == 0:
            return f(0)
        return f(G.dy) * (f(Hy[x, y, z - 1]) - f(Hy[x, y, z])) + f(G.dz) * (f(Hz[x, y, z]) - f(Hz[x, y - 1, z]))
    if tl.polarisation == 'y':
        if x == 0 or z == 0:
            return f(0)
        return f(G.dx) * (f(Hx[x, y, z]) - f(Hx[x, y, z - 1])) + f(G.dz) * (f(Hz[x - 1, y, z]) - f(Hz[x, y, z]))
    if x == 0 or y == 0:
        return f(0)
    return f(G.dx) * (f(Hx[x, y - 1, z]) - f(Hx[x, y, z])) + f(G.dy) * (f(Hy[x, y, z]) - f(Hy[x - 1, y, z]))


def tl_update_magnetic(tl, it, G):
    """gprMax/sources.py:428-451."""
    if it * G.dt >= tl.start and it * G.dt <= tl.stop:
        tl_update_current(tl, it, G)
        tl.current[tl.antpos] = tl_loop_current(tl, G)


def tl_update_electric(tl, it, G):
    """gprMax/sources.py:404-426."""
    if it * G.dt >= tl.start and it * G.dt <= tl.stop:
        tl_update_voltage(tl, it, G)
        d = {'x': G.dx, 'y': G.dy, 'z': G.dz}[tl.polarisation]
        F = getattr(G, 'E' + tl.polarisation)
        F[tl.xcoord, tl.ycoord, tl.zcoord] = -np.float32(tl.voltage[tl.antpos]) / np.float32(d)


def store_outputs(it, G):
    """gprMax/fields_outputs.py:40-67 (field components, then each
    transmission line's antenna voltage and current)."""
    for rx in G.rxs:
        for name, buf in rx.outputs.items():
            buf[it] = getattr(G, name)[rx.xcoord, rx.ycoord, rx.zcoord]
    for tl in G.transmissionlines:
        tl.Vtotal[it] = tl.voltage[tl.antpos]
        tl.Itotal[it] = tl.current[tl.antpos]


def store_snapshot(lib, snap, G):
    out = [np.zeros((snap.nx, snap.ny, snap.nz), dtype=G.floattype) for _ in range(6)]
    lib.funcs['snapshot_fields'](G.nx, G.ny, G.nz, snap.xs, snap.ys, snap.zs, snap.dx, snap.dy, snap.dz,
                                 snap.nx, snap.ny, snap.nz,
                                 ptr(G.Ex), ptr(G.Ey), ptr(G.Ez), ptr(G.Hx), ptr(G.Hy), ptr(G.Hz),
                                 *(ptr(o) for o in out))
    snap.set_from_components(*out)


def initialise_state(G):
    """Zero fields, PML Phi arrays and dispersive T arrays."""
    G.initialise_field_arrays()
    for pml in G.pmls:
        pml.initialise_field_arrays(G.floattype)
    if G.maxpoles > 0:
        G.initialise_dispersive_arrays()


def step(lib, G, it, nthreads):
    """One iteration in the reference's order."""
    f = lib.funcs
    P = G.maxpoles
    store_outputs(it, G)
    for snap in G.snapshots:
        if snap.time == it + 1:
            store_snapshot(lib, snap, G)
    f['update_magnetic'](G.nx, G.ny, G.nz, nthreads, ptr(G.updatecoeffsH), ptr(G.ID),
                         ptr(G.Ex), ptr(G.Ey), ptr(G.Ez), ptr(G.Hx), ptr(G.Hy), ptr(G.Hz))
    for pml in G.pmls:
        _pml_call(lib, G, pml, 1, nthreads)
    for tl in G.transmissionlines:
        tl_update_magnetic(tl, it, G)
    for src in G.magneticdipoles:
        magnetic_dipole_update(src, it, G)
    if P == 0:
        f['update_electric'](G.nx, G.ny, G.nz, nthreads, ptr(G.updatecoeffsE), ptr(G.ID),
                             ptr(G.Ex), ptr(G.Ey), ptr(G.Ez), ptr(G.Hx), ptr(G.Hy), ptr(G.Hz))
    else:
        f['update_electric_dispersive_A'](G.nx, G.ny, G.nz, nthreads, P, ptr(G.updatecoeffsE),
                                          ptr(G.updatecoeffsdispersive), ptr(G.ID),
                                          ptr(G.Tx), ptr(G.Ty), ptr(G.Tz),
                                          ptr(G.Ex), ptr(G.Ey), ptr(G.Ez), ptr(G.Hx), ptr(G.Hy), ptr(G.Hz))
    for pml in G.pmls:
        _pml_call(lib, G, pml, 0, nthreads)
    for src in G.voltagesources:
        voltage_source_update(src, it, G)
    for tl in G.transmissionlines:
        tl_update_electric(tl, it, G)
    for src in G.hertziandipoles:
        hertzian_dipole_update(src, it, G)
    if P > 0:
        f['update_electric_dispersive_B'](G.nx, G.ny, G.nz, nthreads, P, ptr(G.updatecoeffsdispersive),
                                          ptr(G.ID), ptr(G.Tx), ptr(G.Ty), ptr(G.Tz),
                                          ptr(G.Ex), ptr(G.Ey), ptr(G.Ez))


def solve_cpu(currentmodelrun, modelend, G, nthreads=None, iterations=None):
    """Run the CPU solve; returns the solve time in seconds
    (gprMax/model_build_run.py:408-474).  ``iterations`` may cut the run short
    for bounded timing samples."""
    lib = load(G.floattype)
    nthreads = nthreads or G.nthreads or default_threads()
    initialise_state(G)
    n = G.iterations if iterations is None else iterations
    t0 = time.perf_counter()
    for it in range(n):
        step(lib, G, it, nthreads)
    return time.perf_counter() - t0


# ---------------------------------------------------------------------------
# The reference's own compiled kernels (oracle/_ref, built from its .pyx
# sources by oracle/build_ref.py), driven in the same step order.  Used to
# time the reference's CPU path beside the GPU; never as the thing measured.

REF_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref')


def load_reference(flavour='exact'):
    """The compiled reference modules of one build flavour, or None when they
    are not built here."""
    import importlib
    import sys
    tree = os.path.join(REF_DIR, flavour)
    if not os.path.isdir(tree):
        return None
    sys.path.insert(0, tree)
    try:
        mods = {'fields': importlib.import_module('gprMax.fields_updates_ext')}
        for field in ('electric', 'magnetic'):
            for form in ('HORIPML', 'MRIPML'):
                mods[(field, form)] = importlib.import_module(
                    'gprMax.pml_updates.pml_updates_{}_{}_ext'.format(field, form))
    except ImportError:
        return None
    finally:
        sys.path.remove(tree)
    return mods


def _ref_pml(mods, G, pml, field, nthreads):
    """gprMax/pml.py:283-305: the slab's order<n>_<direction> kernel."""
    name = 'electric' if field == 0 else 'magnetic'
    fn = getattr(mods[(name, G.pmlformulation)], 'order{}_{}'.format(len(pml.CFS), pml.direction))
    if field == 0:
        fn(pml.xs, pml.xf, pml.ys, pml.yf, pml.zs, pml.zf, nthreads, G.updatecoeffsE, G.ID,
           G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz, pml.EPhi1, pml.EPhi2, pml.ERA, pml.ERB, pml.ERE, pml.ERF, pml.d)
    else:
        fn(pml.xs, pml.xf, pml.ys, pml.yf, pml.zs, pml.zf, nthreads, G.updatecoeffsH, G.ID,
           G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz, pml.HPhi1, pml.HPhi2, pml.HRA, pml.HRB, pml.HRE, pml.HRF, pml.d)


def step_reference(mods, G, it, nthreads):
    """One iteration of gprMax/model_build_run.py:424-470 with the reference's
    compiled kernels and the NumPy source updates of this module."""
    fu = mods['fields']
    P = G.maxpoles
    store_outputs(it, G)
    fu.update_magnetic(G.nx, G.ny, G.nz, nthreads, G.updatecoeffsH, G.ID, G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz)
    for pml in G.pmls:
        _ref_pml(mods, G, pml, 1, nthreads)
    for tl in G.transmissionlines:
        tl_update_magnetic(tl, it, G)
    for src in G.magneticdipoles:
        magnetic_dipole_update(src, it, G)
    T = (G.Tx, G.Ty, G.Tz) if P else ()
    if P == 0:
        fu.update_electric(G.nx, G.ny, G.nz, nthreads, G.updatecoeffsE, G.ID, G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz)
    elif P == 1:
        fu.update_electric_dispersive_1pole_A(G.nx, G.ny, G.nz, nthreads, G.updatecoeffsE,
                                              G.updatecoeffsdispersive, G.ID, *T,
                                              G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz)
    else:
        fu.update_electric_dispersive_multipole_A(G.nx, G.ny, G.nz, nthreads, P, G.updatecoeffsE,
                                                  G.updatecoeffsdispersive, G.ID, *T,
                                                  G.Ex, G.Ey, G.Ez, G.Hx, G.Hy, G.Hz)
    for pml in G.pmls:
        _ref_pml(mods, G, pml, 0, nthreads)
    for src in G.voltagesources:
        voltage_source_update(src, it, G)
    for tl in G.transmissionlines:
        tl_update_electric(tl, it, G)
    for src in G.hertziandipoles:
        hertzian_dipole_update(src, it, G)
    if P == 1:
        fu.update_electric_dispersive_1pole_B(G.nx, G.ny, G.nz, nthreads, G.updatecoeffsdispersive, G.ID,
                                              *T, G.Ex, G.Ey, G.Ez)
    elif P > 1:
        fu.update_electric_dispersive_multipole_B(G.nx, G.ny, G.nz, nthreads, P, G.updatecoeffsdispersive,
                                                  G.ID, *T, G.Ex, G.Ey, G.Ez)
