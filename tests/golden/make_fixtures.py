"""Convert the reference's golden receiver files into small .npz fixtures.

Reads ``tests/models_basic/<model>/<model>_ref.out`` (and the companion
``<model>.out`` test outputs) from the reference tree with the local HDF5
reader and writes ``tests/golden/<model>[_test].npz`` holding:

* ``traces``  float32 [6, iterations], components Ex, Ey, Ez, Hx, Hy, Hz of rx1
* ``dt``, ``dxdydz``, ``nxyz``, ``iterations``
* ``rx_position``, ``src_position`` (metres), ``src_type``, ``version``

and, from ``tests/other_codes/vs_MoM_MATLAB/antenna_bowtie_fs/antenna_bowtie_fs.out``
(gprMax 3.0.9), ``tests/golden/antenna_bowtie_fs_tl.npz``: the incident voltage
and current of its transmission line with the line's ``dl`` and resistance.

Run once in a container that has the reference mounted:
    python tests/golden/make_fixtures.py /root/reference
"""

import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import h5mini  # noqa: E402

MODELS = ['hertzian_dipole_fs', 'hertzian_dipole_hs', 'hertzian_dipole_dispersive',
          'magnetic_dipole_fs', 'cylinder_Ascan_2D', '2D_ExHyHz', '2D_EyHxHz', '2D_EzHxHy']
ANALYTICAL = 'hertzian_dipole_fs_analytical'
COMPONENTS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def convert(path, out):
    d = h5mini.read_file(path)
    root = d['/']['attrs']
    traces = np.stack([d['/rxs/rx1/' + c]['data'] for c in COMPONENTS]).astype(np.float32)
    dxdydz = root.get('dx_dy_dz', root.get('dx, dy, dz'))
    nxyz = root.get('nx_ny_nz', root.get('nx, ny, nz'))
    np.savez_compressed(out, traces=traces, dt=np.float64(root['dt']),
                        dxdydz=np.asarray(dxdydz, dtype=np.float64),
                        nxyz=np.asarray(nxyz, dtype=np.int64),
                        iterations=np.int64(root['Iterations']),
                        rx_position=np.asarray(d['/rxs/rx1']['attrs']['Position'], dtype=np.float64),
                        src_position=np.asarray(d['/srcs/src1']['attrs']['Position'], dtype=np.float64),
                        src_type=np.array(d['/srcs/src1']['attrs'].get('Type', '')),
                        version=np.array(root.get('gprMax', '')))


def convert_tl(path, out):
    """The incident voltage and current of the first transmission line of a
    reference output file (``/tls/tl1/Vinc``, ``Iinc``) with the line's and
    the run's attributes."""
    d = h5mini.read_file(path)
    root = d['/']['attrs']
    tl = d['/tls/tl1']['attrs']
    np.savez_compressed(out, Vinc=np.asarray(d['/tls/tl1/Vinc']['data'], dtype=np.float32),
                        Iinc=np.asarray(d['/tls/tl1/Iinc']['data'], dtype=np.float32),
                        dt=np.float64(root['dt']), iterations=np.int64(root['Iterations']),
                        dxdydz=np.asarray(root.get('dx_dy_dz', root.get('dx, dy, dz')), dtype=np.float64),
                        dl=np.float64(tl['dl']), resistance=np.float64(tl['Resistance']),
                        position=np.asarray(tl['Position'], dtype=np.float64),
                        version=np.array(root.get('gprMax', '')))


def main(reference):
    here = os.path.dirname(os.path.abspath(__file__))
    base = os.path.join(reference, 'tests', 'models_basic')
    for model in MODELS:
        convert(os.path.join(base, model, model + '_ref.out'), os.path.join(here, model + '_ref.npz'))
        convert(os.path.join(base, model, model + '.out'), os.path.join(here, model + '_test.npz'))
    convert(os.path.join(base, ANALYTICAL, ANALYTICAL + '.out'), os.path.join(here, ANALYTICAL + '_test.npz'))
    bowtie = os.path.join(reference, 'tests', 'other_codes', 'vs_MoM_MATLAB', 'antenna_bowtie_fs')
    convert_tl(os.path.join(bowtie, 'antenna_bowtie_fs.out'), os.path.join(here, 'antenna_bowtie_fs_tl.npz'))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
