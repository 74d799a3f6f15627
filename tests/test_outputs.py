"""Receiver output files (gprmax_amd.outputs, gprmax_amd.hdf5).

The writer is checked against a file written by libhdf5 itself (the
reference's tests/models_basic/2D_EzHxHy/2D_EzHxHy_ref.out, kept as a
fixture): the same tree written here carries the same object header
messages, byte for byte except file addresses and global-heap indices."""

import os

import numpy as np
import pytest

from gprmax_amd import hdf5
from gprmax_amd.outputs import get_output_data, merge_files, write_hdf5_outputfile

HERE = os.path.dirname(os.path.abspath(__file__))
LIBHDF5_FILE = os.path.join(HERE, 'golden', '2D_EzHxHy_ref.out')


def _messages(reader, addr):
    """(type, flags, body) of each message, file addresses blanked."""
    out = []
    for mtype, pos, size in reader.messages(addr):
        if mtype in (0x00, 0x10, 0x12):       # NIL, continuation, modification time
            continue
        body = bytearray(reader.buf[pos:pos + size])
        if mtype == 0x11:                     # symbol table: B-tree and heap addresses
            body[:] = bytes(len(body))
        elif mtype == 0x08:                   # layout v3 contiguous: data address
            body[2:10] = bytes(8)
        elif mtype == 0x0C and body[8 + ((body[2] | body[3] << 8) + 7) // 8 * 8] & 0x0F == 9:
            body[-12:] = bytes(12)            # vlen string: heap address and index
        out.append((mtype, reader.buf[pos - 4], bytes(body)))
    return out


def _tree_from_file(reader):
    root = hdf5.Group()

    def build(entry, grp):
        for k, v in reader.attributes(entry['header']).items():
            grp.attrs[k] = v
        for name, child in sorted(reader.group_members(entry).items()):
            if child.get('cache') == 1:
                build(child, grp.group(name))
            else:
                grp.dataset(name, reader.dataset(child['header']))
    build(reader.root, root)
    return root


def test_writer_matches_libhdf5_layout(tmp_path):
    ref = hdf5._Reader(LIBHDF5_FILE)
    path = str(tmp_path / 'copy.out')
    hdf5.write_file(path, _tree_from_file(ref))
    ours = hdf5._Reader(path)
    assert ours.buf[:24] == ref.buf[:24]      # superblock v0, 8-byte offsets, K = 4 / 16
    count = 0

    def walk(a, b, where):
        nonlocal count
        assert _messages(ref, a['header']) == _messages(ours, b['header']), where
        count += 1
        if a.get('cache') == 1:
            ma, mb = ref.group_members(a), ours.group_members(b)
            assert list(ma) == list(mb), where
            for k in ma:
                walk(ma[k], mb[k], where + '/' + k)
    walk(ref.root, ours.root, '')
    assert count == 11                        # root, srcs, src1, rxs, rx1 and six datasets
    a, b = hdf5.read_file(LIBHDF5_FILE), hdf5.read_file(path)
    assert list(a) == list(b)
    for k in a:
        assert a[k]['attrs'].keys() == b[k]['attrs'].keys()
        for n, v in a[k]['attrs'].items():
            assert np.array_equal(np.asarray(v), np.asarray(b[k]['attrs'][n])), (k, n)
        if 'data' in a[k]:
            assert a[k]['data'].dtype == b[k]['data'].dtype and np.array_equal(a[k]['data'], b[k]['data'])


def test_fixture_traces_match_the_npz_fixture():
    d = hdf5.read_file(LIBHDF5_FILE)
    npz = np.load(os.path.join(HERE, 'golden', '2D_EzHxHy_ref.npz'))
    for c, name in enumerate(('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')):
        assert np.array_equal(d['/rxs/rx1/' + name]['data'], npz['traces'][c])


def _small_model(n, rx_shift=0):
    from gprmax_amd.grid import ModelBuilder
    b = ModelBuilder(12, 11, 10, 1e-3, 1e-3, 1e-3, 30, pml_cells=3)
    b.G.title = 'trace {}'.format(n)
    b.add_waveform('ricker', 1, 40e9, 'w')
    b.add_hertzian_dipole('z', (5 + n, 5, 5), 'w')
    b.add_magnetic_dipole('x', (4, 6, 5), 'w')
    b.add_rx((7 + n, 5, 5))
    b.add_rx((6, 7 + rx_shift, 4), name='second')
    G = b.finalise()
    rng = np.random.default_rng(n)
    for rx in G.rxs:
        for c in rx.outputs:
            rx.outputs[c][:] = rng.standard_normal(G.iterations).astype(np.float32)
    return G


def test_output_file_contents(tmp_path):
    G = _small_model(0)
    path = str(tmp_path / 'model.out')
    write_hdf5_outputfile(path, G)
    d = hdf5.read_file(path)
    root = d['/']['attrs']
    assert root['gprMax'] and root['Title'] == 'trace 0' and root['Iterations'] == G.iterations == 30
    assert list(root['nx_ny_nz']) == [12, 11, 10] and np.allclose(root['dx_dy_dz'], 1e-3)
    assert root['dt'] == G.dt and root['nsrc'] == 2 and root['nrx'] == 2
    assert list(root['srcsteps']) == [0, 0, 0] and list(root['rxsteps']) == [0, 0, 0]
    # sources in the order voltage, Hertzian, magnetic
    assert d['/srcs/src1']['attrs']['Type'] == 'HertzianDipole'
    assert d['/srcs/src2']['attrs']['Type'] == 'MagneticDipole'
    assert np.allclose(d['/srcs/src2']['attrs']['Position'], (4e-3, 6e-3, 5e-3))
    assert d['/rxs/rx2']['attrs']['Name'] == 'second'
    for n, rx in enumerate(G.rxs):
        for c, values in rx.outputs.items():
            got = d['/rxs/rx{}/{}'.format(n + 1, c)]['data']
            assert got.dtype == np.float32 and np.array_equal(got, values)
    data, dt = get_output_data(path, 2, 'Hy')
    assert np.array_equal(data, G.rxs[1].outputs['Hy']) and dt == G.dt
    with pytest.raises(Exception, match='available output'):
        get_output_data(path, 1, 'Ix')


def test_merge_bscan_files(tmp_path):
    base = str(tmp_path / 'bscan')
    models = [_small_model(n) for n in range(3)]
    for n, G in enumerate(models):
        write_hdf5_outputfile(base + str(n + 1) + '.out', G)
    out = merge_files(base, removefiles=True)
    assert out == base + '_merged.out'
    assert sorted(os.listdir(str(tmp_path))) == ['bscan_merged.out']
    d = hdf5.read_file(out)
    root = d['/']['attrs']
    assert root['Title'] == 'trace 0' and root['Iterations'] == 30 and root['nrx'] == 2
    for r in range(2):
        for c in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'):
            got = d['/rxs/rx{}/{}'.format(r + 1, c)]['data']
            assert got.shape == (30, 3) and got.dtype == np.float32
            for n, G in enumerate(models):
                assert np.array_equal(got[:, n], G.rxs[r].outputs[c])
    data, _ = get_output_data(out, 1, 'Ez')
    assert data.shape == (30, 3)


def test_groups_with_many_members(tmp_path):
    """More members than one symbol-table node holds (8): several nodes
    under one B-tree node, keys in name order."""
    g = hdf5.Group()
    names = ['rx{}'.format(n) for n in range(1, 40)]
    for n, name in enumerate(names):
        g.dataset('/rxs/' + name, np.full(3, n, dtype=np.float64))
    path = str(tmp_path / 'many.out')
    hdf5.write_file(path, g)
    d = hdf5.read_file(path)
    for n, name in enumerate(names):
        assert np.array_equal(d['/rxs/' + name]['data'], np.full(3, n))


def test_snapshot_vti_file(tmp_path):
    """A .vti snapshot: the reference's XML header, then per field a UInt32
    byte count and the cell data, three components per cell, x fastest."""
    import re
    import struct
    from gprmax_amd.grid import ModelBuilder
    b = ModelBuilder(12, 10, 8, 1e-3, 2e-3, 1e-3, 10, pml_cells=0)
    snap = b.add_snapshot(2, 0, 2, 10, 10, 8, 2, 2, 3, 5, filename='snap')
    rng = np.random.default_rng(9)
    comps = [rng.standard_normal((snap.nx, snap.ny, snap.nz)).astype(np.float32) for _ in range(6)]
    snap.set_from_components(*comps)
    snap.filename = str(tmp_path / 'snap.vti')
    snap.write_vtk_imagedata(None, b.G)
    raw = open(snap.filename, 'rb').read()
    head, _, rest = raw.partition(b'<AppendedData encoding="raw">\n_')
    text = head.decode()
    assert '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">' in text
    # the extent mixes start cells with end cells divided by the step, as the
    # reference writes it (gprMax/snapshots.py:145-146)
    assert 'WholeExtent="2 5 0 5 2 3"' in text and 'Spacing="0.002 0.004 0.003"' in text
    n = snap.nx * snap.ny * snap.nz
    assert (snap.nx, snap.ny, snap.nz) == (4, 5, 2)
    offsets = [int(v) for v in re.findall(r'offset="(\d+)"', text)]
    assert offsets == [0, 4 + 12 * n]
    for f, off in enumerate(offsets):
        size, = struct.unpack('I', rest[off:off + 4])
        assert size == 12 * n
        data = np.frombuffer(rest, dtype=np.float32, count=3 * n, offset=off + 4)
        for c in range(3):
            for (i, j, k) in ((0, 0, 0), (3, 1, 1), (2, 4, 0)):
                assert data[c + 3 * (i + snap.nx * (j + snap.ny * k))] == comps[3 * f + c][i, j, k]
    assert rest.endswith(b'\n</AppendedData>\n</VTKFile>')
    assert len(rest) == 2 * (4 + 12 * n) + len(b'\n</AppendedData>\n</VTKFile>')


def test_transmission_line_groups(tmp_path):
    """/tls/tl<n> per line (gprMax/fields_outputs.py:139-151): Position,
    Resistance and dl attributes, Vinc, Iinc, Vtotal, Itotal datasets."""
    from oracle import solver as osolver
    from tl_models import tl_model
    G = tl_model(iterations=40)
    osolver.solve_cpu(1, 1, G, nthreads=4)
    path = str(tmp_path / 'tl.out')
    write_hdf5_outputfile(path, G)
    d = hdf5.read_file(path)
    assert d['/']['attrs']['nsrc'] == 3 + 4
    for n, tl in enumerate(G.transmissionlines):
        g = d['/tls/tl{}'.format(n + 1)]
        assert np.allclose(g['attrs']['Position'], (tl.xcoord * 1e-3, tl.ycoord * 1e-3, tl.zcoord * 1e-3))
        assert g['attrs']['Resistance'] == tl.resistance and g['attrs']['dl'] == tl.dl
        for name in ('Vinc', 'Iinc', 'Vtotal', 'Itotal'):
            got = d['/tls/tl{}/{}'.format(n + 1, name)]['data']
            assert got.dtype == np.float32 and np.array_equal(got, getattr(tl, name))
