"""Receiver output files: the ``.out`` file of one model run
(gprMax/fields_outputs.py:111-161) and the merge of a B-scan's per-trace
files into one (tools/outputfiles_merge.py:29-117), in HDF5 through
gprmax_amd.hdf5 (no h5py needed).

The file layout is the reference's: root attributes gprMax, Title,
Iterations, nx_ny_nz, dx_dy_dz, dt, nsrc, nrx, srcsteps, rxsteps; a group
/srcs/src<n> (attributes Type, Position) per voltage source, Hertzian dipole
and magnetic dipole, in that order; a group /tls/tl<n> (attributes Position,
Resistance, dl; datasets Vinc, Iinc, Vtotal, Itotal) per transmission line;
a group /rxs/rx<n> (attributes Name, Position) per receiver holding one
dataset per output component.
"""

import glob
import os

import numpy as np

from .hdf5 import Group, read_file, write_file
from .utilities import GeneralError

# The gprMax release whose output format this follows (gprMax/_version.py);
# gprMax itself passes its own version.
GPRMAX_VERSION = '3.1.7'


def write_hdf5_outputfile(outputfile, G, version=GPRMAX_VERSION):
    """Write the receiver and transmission-line outputs of grid G
    (gprMax/fields_outputs.py:111-161)."""
    f = Group()
    f.attrs['gprMax'] = version
    f.attrs['Title'] = G.title
    f.attrs['Iterations'] = int(G.iterations)
    f.attrs['nx_ny_nz'] = (int(G.nx), int(G.ny), int(G.nz))
    f.attrs['dx_dy_dz'] = (float(G.dx), float(G.dy), float(G.dz))
    f.attrs['dt'] = float(G.dt)
    srclist = list(G.voltagesources) + list(G.hertziandipoles) + list(G.magneticdipoles)
    f.attrs['nsrc'] = len(srclist) + len(getattr(G, 'transmissionlines', []))
    f.attrs['nrx'] = len(G.rxs)
    f.attrs['srcsteps'] = [int(v) for v in getattr(G, 'srcsteps', (0, 0, 0))]
    f.attrs['rxsteps'] = [int(v) for v in getattr(G, 'rxsteps', (0, 0, 0))]
    for n, src in enumerate(srclist):
        grp = f.group('/srcs/src{}'.format(n + 1))
        grp.attrs['Type'] = type(src).__name__
        grp.attrs['Position'] = (src.xcoord * G.dx, src.ycoord * G.dy, src.zcoord * G.dz)
    for n, tl in enumerate(getattr(G, 'transmissionlines', [])):
        path = '/tls/tl{}'.format(n + 1)
        grp = f.group(path)
        grp.attrs['Position'] = (tl.xcoord * G.dx, tl.ycoord * G.dy, tl.zcoord * G.dz)
        grp.attrs['Resistance'] = float(tl.resistance)
        grp.attrs['dl'] = float(tl.dl)
        for name in ('Vinc', 'Iinc', 'Vtotal', 'Itotal'):
            f.dataset(path + '/' + name, getattr(tl, name))
    for n, rx in enumerate(G.rxs):
        path = '/rxs/rx{}'.format(n + 1)
        grp = f.group(path)
        if rx.ID:
            grp.attrs['Name'] = rx.ID
        grp.attrs['Position'] = (rx.xcoord * G.dx, rx.ycoord * G.dy, rx.zcoord * G.dz)
        for output, values in rx.outputs.items():
            f.dataset(path + '/' + output, values)
    write_file(outputfile, f)


def get_output_data(filename, rxnumber, rxcomponent):
    """(data, dt) of one receiver component: an A-scan [iterations], or a
    B-scan [iterations, traces] from a merged file
    (tools/outputfiles_merge.py:29-62)."""
    d = read_file(filename)
    root = d['/']['attrs']
    if int(root['nrx']) == 0:
        raise GeneralError('No receivers found in {}'.format(filename))
    path = '/rxs/rx{}'.format(rxnumber)
    available = [k.rsplit('/', 1)[1] for k in d if k.startswith(path + '/')]
    if rxcomponent not in available:
        raise GeneralError('{} output requested to plot, but the available output for receiver 1 is {}'.format(
            rxcomponent, ', '.join(available)))
    return d[path + '/' + rxcomponent]['data'], root['dt']


def merge_files(basefilename, removefiles=False, version=GPRMAX_VERSION):
    """Merge the traces of <basefilename>1.out, 2.out, ... into
    <basefilename>_merged.out, one [iterations, traces] dataset per receiver
    component (tools/outputfiles_merge.py:65-117)."""
    outputfile = basefilename + '_merged.out'
    files = [f for f in glob.glob(basefilename + '[0-9]*.out') if '_merged' not in f]
    modelruns = len(files)
    if modelruns == 0:
        raise GeneralError('no output files match {}'.format(basefilename + '[0-9]*.out'))
    out = Group()
    merged = {}
    for model in range(modelruns):
        d = read_file(basefilename + str(model + 1) + '.out')
        attrs = d['/']['attrs']
        nrx = int(attrs['nrx'])
        if model == 0:
            out.attrs['Title'] = attrs['Title']
            out.attrs['gprMax'] = version
            out.attrs['Iterations'] = attrs['Iterations']
            out.attrs['dt'] = attrs['dt']
            out.attrs['nrx'] = attrs['nrx']
            for rx in range(1, nrx + 1):
                path = '/rxs/rx{}'.format(rx)
                out.group(path)
                for key in d:
                    if key.startswith(path + '/'):
                        merged[key] = np.zeros((int(attrs['Iterations']), modelruns), dtype=d[key]['data'].dtype)
        for rx in range(1, nrx + 1):
            path = '/rxs/rx{}'.format(rx)
            for key in d:
                if key.startswith(path + '/'):
                    merged[key][:, model] = d[key]['data']
    for key, data in merged.items():
        out.dataset(key, data)
    write_file(outputfile, out)
    if removefiles:
        for model in range(modelruns):
            os.remove(basefilename + str(model + 1) + '.out')
    return outputfile
