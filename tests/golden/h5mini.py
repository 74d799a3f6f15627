"""Minimal read-only HDF5 parser for the reference's receiver output files.

Covers what those files use: superblock version 0, version-1 object headers
(with continuation blocks), old-style groups (symbol-table B-trees, local
heaps), contiguous and compact datasets of fixed-size numbers, and attributes
holding numbers or variable-length strings (global heap).  Written from the
published HDF5 file-format specification; h5py is not available here.
"""

import struct

import numpy as np


class H5Error(Exception):
    pass


class _Reader:
    def __init__(self, path):
        with open(path, 'rb') as fh:
            self.buf = fh.read()
        sig = b'\x89HDF\r\n\x1a\n'
        if self.buf[:8] != sig:
            raise H5Error('not an HDF5 file')
        version = self.buf[8]
        if version != 0:
            raise H5Error('superblock version {} not supported'.format(version))
        self.so = self.buf[13]  # size of offsets
        self.sl = self.buf[14]  # size of lengths
        pos = 24
        self.base = self._off(pos)
        pos += 4 * self.so  # base, free-space, EOF, driver info
        self.root = self._symbol_entry(pos)
        self._gheap = {}

    # -- primitives -------------------------------------------------------
    def _u(self, pos, n):
        return int.from_bytes(self.buf[pos:pos + n], 'little')

    def _off(self, pos):
        return self._u(pos, self.so)

    def _len(self, pos):
        return self._u(pos, self.sl)

    def _symbol_entry(self, pos):
        name_off = self._off(pos)
        header = self._off(pos + self.so)
        cache = self._u(pos + 2 * self.so, 4)
        scratch = pos + 2 * self.so + 8
        entry = {'name_off': name_off, 'header': header, 'cache': cache}
        if cache == 1:
            entry['btree'] = self._off(scratch)
            entry['heap'] = self._off(scratch + self.so)
        return entry

    # -- object headers ----------------------------------------------------
    def messages(self, addr):
        if self.buf[addr] != 1:
            raise H5Error('object header version {} not supported'.format(self.buf[addr]))
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            pos, end = start, start + size
            while pos + 8 <= end and len(out) < nmsg:
                mtype = self._u(pos, 2)
                msize = self._u(pos + 2, 2)
                body = pos + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self._off(body), self._len(body + self.so)))
                out.append((mtype, body, msize))
                pos = body + msize
        return out

    # -- messages ------------------------------------------------------------
    def dataspace(self, pos):
        version = self.buf[pos]
        ndims = self.buf[pos + 1]
        if version == 1:
            p = pos + 8
        elif version == 2:
            p = pos + 4
        else:
            raise H5Error('dataspace version {}'.format(version))
        return tuple(self._len(p + n * self.sl) for n in range(ndims))

    def datatype(self, pos):
        cls = self.buf[pos] & 0x0F
        bits = self._u(pos + 1, 3)
        size = self._u(pos + 4, 4)
        if cls == 0:
            signed = bool(bits & 0x08)
            return np.dtype(('<i' if signed else '<u') + str(size))
        if cls == 1:
            return np.dtype('<f' + str(size))
        if cls == 3:
            return np.dtype('S' + str(size))
        if cls == 9:
            return 'vlen-string' if (bits & 0x0F) == 1 else 'vlen'
        raise H5Error('datatype class {} not supported'.format(cls))

    def _vlen_string(self, pos):
        length = self._u(pos, 4)
        coll = self._off(pos + 4)
        index = self._u(pos + 4 + self.so, 4)
        if coll not in self._gheap:
            self._gheap[coll] = self._read_gcol(coll)
        data = self._gheap[coll].get(index, b'')
        return data[:length].decode('utf-8', 'replace')

    def _read_gcol(self, addr):
        if self.buf[addr:addr + 4] != b'GCOL':
            raise H5Error('bad global heap')
        size = self._len(addr + 8)
        pos, end = addr + 8 + self.sl, addr + size
        objects = {}
        while pos + 8 + self.sl <= end:
            index = self._u(pos, 2)
            osize = self._len(pos + 8)
            if index == 0:
                break
            data = self.buf[pos + 8 + self.sl:pos + 8 + self.sl + osize]
            objects[index] = data
            pos += 8 + self.sl + ((osize + 7) // 8) * 8
        return objects

    def attributes(self, addr):
        attrs = {}
        for mtype, pos, size in self.messages(addr):
            if mtype != 0x0C:
                continue
            version = self.buf[pos]
            nlen = self._u(pos + 2, 2)
            tlen = self._u(pos + 4, 2)
            slen = self._u(pos + 6, 2)
            p = pos + 8
            pad = (lambda n: ((n + 7) // 8) * 8) if version == 1 else (lambda n: n)
            if version == 3:
                p += 1
            name = self.buf[p:p + nlen].split(b'\x00')[0].decode()
            p += pad(nlen)
            dtype = self.datatype(p)
            p += pad(tlen)
            shape = self.dataspace(p)
            p += pad(slen)
            count = int(np.prod(shape)) if shape else 1
            if dtype == 'vlen-string':
                vals = [self._vlen_string(p + n * (4 + self.so + 4)) for n in range(count)]
                attrs[name] = vals[0] if not shape else vals
            elif dtype == 'vlen':
                attrs[name] = None
            else:
                arr = np.frombuffer(self.buf, dtype=dtype, count=count, offset=p).copy()
                attrs[name] = arr.reshape(shape) if shape else arr[0]
        return attrs

    def dataset(self, addr):
        shape = dtype = layout = None
        for mtype, pos, size in self.messages(addr):
            if mtype == 0x01:
                shape = self.dataspace(pos)
            elif mtype == 0x03:
                dtype = self.datatype(pos)
            elif mtype == 0x08:
                layout = pos
        if shape is None or dtype is None or layout is None:
            raise H5Error('object at {} is not a simple dataset'.format(addr))
        version = self.buf[layout]
        count = int(np.prod(shape))
        if version == 3:
            cls = self.buf[layout + 1]
            if cls == 1:
                data_addr = self._off(layout + 2)
            elif cls == 0:
                data_addr = layout + 4
            else:
                raise H5Error('chunked datasets not supported')
        elif version in (1, 2):
            cls = self.buf[layout + 2]
            if cls != 1:
                raise H5Error('layout class {} not supported'.format(cls))
            data_addr = self._off(layout + 8)
        else:
            raise H5Error('layout version {}'.format(version))
        if data_addr == (1 << (8 * self.so)) - 1:
            return np.zeros(shape, dtype=dtype)
        return np.frombuffer(self.buf, dtype=dtype, count=count, offset=data_addr).reshape(shape).copy()

    # -- groups --------------------------------------------------------------
    def group_members(self, entry):
        if entry.get('cache') != 1:
            for mtype, pos, size in self.messages(entry['header']):
                if mtype == 0x11:
                    entry = dict(entry, btree=self._off(pos), heap=self._off(pos + self.so), cache=1)
                    break
            else:
                raise H5Error('object is not a group')
        heap = entry['heap']
        if self.buf[heap:heap + 4] != b'HEAP':
            raise H5Error('bad local heap')
        data = self._off(heap + 8 + 2 * self.sl)
        members = {}
        self._walk(entry['btree'], data, members)
        return members

    def _walk(self, node, heapdata, members):
        if self.buf[node:node + 4] == b'SNOD':
            n = self._u(node + 6, 2)
            pos = node + 8
            size = 2 * self.so + 24
            for k in range(n):
                e = self._symbol_entry(pos + k * size)
                name_pos = heapdata + e['name_off']
                name = self.buf[name_pos:self.buf.index(b'\x00', name_pos)].decode()
                members[name] = e
            return
        if self.buf[node:node + 4] != b'TREE':
            raise H5Error('bad B-tree node')
        used = self._u(node + 6, 2)
        pos = node + 8 + 2 * self.so
        for k in range(used):
            child = self._off(pos + self.sl + k * (self.sl + self.so))
            self._walk(child, heapdata, members)


def read_file(path):
    """Return ``{'attrs': {...}, '<group path>': {...}}`` for the whole file.

    Each entry maps an object path (e.g. '/rxs/rx1/Ez') to either a dict with
    'attrs' (groups) or a dict with 'data' and 'attrs' (datasets).
    """
    r = _Reader(path)
    out = {}

    def visit(entry, path):
        attrs = r.attributes(entry['header'])
        is_group = entry.get('cache') == 1 or any(t == 0x11 for t, _, _ in r.messages(entry['header']))
        if is_group:
            out[path or '/'] = {'attrs': attrs}
            for name, child in sorted(r.group_members(entry).items()):
                visit(child, path + '/' + name)
        else:
            out[path] = {'attrs': attrs, 'data': r.dataset(entry['header'])}

    visit(r.root, '')
    return out
