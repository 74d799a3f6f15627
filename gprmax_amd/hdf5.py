"""Minimal HDF5 files without libhdf5: the receiver output files of the solve
path (``.out``, gprMax/fields_outputs.py:111-161) and merged B-scans
(tools/outputfiles_merge.py:65-117).

Written from the published HDF5 file-format specification, in the layout
libhdf5 itself writes with its default (earliest) format bounds, the layout of
the reference's own ``.out`` files:

* superblock version 0, 8-byte offsets and lengths, group leaf/internal node
  K = 4/16;
* version-1 object headers;
* old-style groups: symbol-table message, version-1 B-tree (type 0) over
  symbol-table nodes, names in a local heap;
* contiguous datasets (dataspace v1 with maximum dimensions, fill value v2,
  layout v3) of little-endian integers and IEEE floats;
* version-1 attributes of numbers, 1-D number arrays and variable-length
  UTF-8 strings (kept in one global heap collection).

The reader covers the same structures plus continuation blocks, compact
datasets and the older layout versions found in files written by libhdf5.
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


# ---------------------------------------------------------------------------
# Writer
# ---------------------------------------------------------------------------
UNDEF = 0xFFFFFFFFFFFFFFFF
_LEAF_K, _NODE_K = 4, 16                       # superblock group K values
_SNOD_SIZE = 8 + 2 * _LEAF_K * 40              # symbol-table node, full capacity
_TREE_SIZE = 8 + 2 * 8 + (2 * _NODE_K + 1) * 8 + 2 * _NODE_K * 8
_GCOL_MIN = 4096                               # libhdf5 reads a collection in 4 KiB


class Group:
    """A group to write: ordered attributes and members (groups or datasets)."""

    def __init__(self):
        self.attrs = {}
        self.members = {}

    def group(self, path):
        """The subgroup at a '/'-separated path, created as needed."""
        g = self
        for name in [p for p in path.split('/') if p]:
            if name not in g.members:
                g.members[name] = Group()
            g = g.members[name]
            if not isinstance(g, Group):
                raise H5Error('{} is a dataset'.format(name))
        return g

    def dataset(self, path, data):
        head, _, name = path.rstrip('/').rpartition('/')
        arr = np.ascontiguousarray(data)
        if arr.dtype.kind not in 'iuf':
            raise H5Error('dataset {}: only integer and float data'.format(path))
        self.group(head).members[name] = arr
        return arr


def _pad8(n):
    return (n + 7) // 8 * 8


def _dtype_message(dt):
    """Datatype message body (version 1) for an integer or IEEE float."""
    dt = np.dtype(dt).newbyteorder('<')
    size = dt.itemsize
    if dt.kind in 'iu':
        bits = 0x08 if dt.kind == 'i' else 0x00
        return struct.pack('<B3sI', 0x10, bytes([bits, 0, 0]), size) + struct.pack('<HH', 0, 8 * size)
    if dt.kind == 'f' and size in (4, 8):
        # byte 1: mantissa normalisation 2 (implied msb); byte 2: sign bit
        sign, eloc, esize, msize, bias = (31, 23, 8, 23, 127) if size == 4 else (63, 52, 11, 52, 1023)
        return (struct.pack('<B3sI', 0x11, bytes([0x20, sign, 0]), size)
                + struct.pack('<HHBBBBI', 0, 8 * size, eloc, esize, 0, msize, bias))
    raise H5Error('unsupported data type {}'.format(dt))


def _vlen_string_type():
    # class 9 (variable length), type 1 (string), null-terminated, UTF-8;
    # base type: 1-byte unsigned integer
    return struct.pack('<B3sI', 0x19, bytes([0x01, 0x01, 0]), 16) + struct.pack('<B3sIHH', 0x10, b'\0\0\0', 1, 0, 8)


def _dataspace_message(shape):
    """Dataspace message body (version 1), maximum dimensions = dimensions."""
    if not shape:
        return struct.pack('<BBBB4x', 1, 0, 0, 0)
    body = struct.pack('<BBBB4x', 1, len(shape), 1, 0)
    body += b''.join(struct.pack('<Q', n) for n in shape) * 2
    return body


class _Out:
    def __init__(self):
        self.buf = bytearray()

    def alloc(self, size):
        addr = len(self.buf)
        self.buf += bytes(_pad8(size))
        return addr

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data


def _message(mtype, body, flags=0):
    body = bytes(body)
    return struct.pack('<HHB3x', mtype, _pad8(len(body)), flags) + body + bytes(_pad8(len(body)) - len(body))


def _object_header(messages):
    data = b''.join(messages)
    return struct.pack('<BBHII4x', 1, 0, len(messages), 1, len(data)) + data


class _Writer:
    def __init__(self, root):
        self.root = root
        self.out = _Out()
        self.strings = []

    # attributes ------------------------------------------------------------
    def _collect_strings(self, g):
        for v in g.attrs.values():
            if isinstance(v, str):
                self.strings.append(v.encode('utf-8'))
        for m in g.members.values():
            if isinstance(m, Group):
                self._collect_strings(m)

    def _attribute(self, name, value):
        nm = name.encode('utf-8') + b'\0'
        if isinstance(value, str):
            tbody, sbody = _vlen_string_type(), _dataspace_message(())
            raw = value.encode('utf-8')
            index = self.string_index[raw]
            data = struct.pack('<IQI', len(raw), self.gcol, index)
        else:
            arr = np.asarray(value)
            if arr.dtype.kind == 'b':
                arr = arr.astype(np.int8)
            if arr.dtype.kind == 'i' and not isinstance(value, np.ndarray) and not isinstance(value, np.generic):
                arr = arr.astype(np.int64)
            if arr.dtype.kind == 'f' and not isinstance(value, np.ndarray) and not isinstance(value, np.generic):
                arr = arr.astype(np.float64)
            if arr.ndim > 1:
                raise H5Error('attribute {}: at most one dimension'.format(name))
            arr = arr.astype(arr.dtype.newbyteorder('<'))
            tbody, sbody = _dtype_message(arr.dtype), _dataspace_message(arr.shape)
            data = arr.tobytes()
        body = struct.pack('<BBHHH', 1, 0, len(nm), len(tbody), len(sbody))
        body += nm + bytes(_pad8(len(nm)) - len(nm))
        body += tbody + bytes(_pad8(len(tbody)) - len(tbody))
        body += sbody + bytes(_pad8(len(sbody)) - len(sbody))
        body += data
        return _message(0x0C, body, flags=4)   # libhdf5 marks attributes "do not share"

    # objects ------------------------------------------------------------------
    def _dataset(self, arr):
        out = self.out
        data_addr = out.alloc(arr.nbytes) if arr.nbytes else UNDEF
        if arr.nbytes:
            out.put(data_addr, arr.astype(arr.dtype.newbyteorder('<')).tobytes())
        msgs = [_message(0x01, _dataspace_message(arr.shape)),
                _message(0x03, _dtype_message(arr.dtype), flags=1),
                _message(0x05, struct.pack('<BBBBI', 2, 2, 2, 1, 0), flags=1),
                _message(0x08, struct.pack('<BBQQ', 3, 1, data_addr, arr.nbytes), flags=1)]
        hdr = _object_header(msgs)
        addr = out.alloc(len(hdr))
        out.put(addr, hdr)
        return addr

    def _group(self, g):
        """Writes the group's members, name heap, B-tree and header; returns
        (header, B-tree, heap) addresses."""
        entries = []
        for name, m in g.members.items():
            if isinstance(m, Group):
                h, bt, hp = self._group(m)
                entries.append((name.encode('utf-8'), h, 1, bt, hp))
            else:
                entries.append((name.encode('utf-8'), self._dataset(m), 0, 0, 0))
        entries.sort(key=lambda e: e[0])
        out = self.out
        # local heap: "" at offset 0, then the names
        heapdata = bytearray(8)
        offsets = []
        for e in entries:
            offsets.append(len(heapdata))
            heapdata += e[0] + b'\0'
            heapdata += bytes(_pad8(len(heapdata)) - len(heapdata))
        heap = out.alloc(32 + len(heapdata))
        out.put(heap, b'HEAP' + struct.pack('<B3xQQQ', 0, len(heapdata), 1, heap + 32) + bytes(heapdata))
        # symbol-table nodes of at most 2K entries, under one B-tree node
        per = 2 * _LEAF_K
        chunks = [list(range(n, min(n + per, len(entries)))) for n in range(0, len(entries), per)] or [[]]
        if len(chunks) > 2 * _NODE_K:
            raise H5Error('groups of more than {} members are not supported'.format(per * 2 * _NODE_K))
        tree = out.alloc(_TREE_SIZE)
        keys, children = [0], []
        for chunk in chunks:
            node = out.alloc(_SNOD_SIZE)
            body = bytearray(b'SNOD' + struct.pack('<BBH', 1, 0, len(chunk)))
            for n in chunk:
                name, h, cache, bt, hp = entries[n]
                body += struct.pack('<QQII', offsets[n], h, cache, 0)
                body += struct.pack('<QQ', bt, hp) if cache == 1 else bytes(16)
            out.put(node, bytes(body))
            children.append(node)
            keys.append(offsets[chunk[-1]] if chunk else 0)
        if not entries:
            children, keys = [], [0]
        body = bytearray(b'TREE' + struct.pack('<BBH', 0, 0, len(children)) + struct.pack('<QQ', UNDEF, UNDEF))
        for n, child in enumerate(children):
            body += struct.pack('<QQ', keys[n], child)
        body += struct.pack('<Q', keys[len(children)])
        out.put(tree, bytes(body))
        msgs = [_message(0x11, struct.pack('<QQ', tree, heap))]
        msgs += [self._attribute(k, v) for k, v in g.attrs.items()]
        hdr = _object_header(msgs)
        addr = out.alloc(len(hdr))
        out.put(addr, hdr)
        return addr, tree, heap

    def write(self, path):
        out = self.out
        out.alloc(96)                                   # superblock + root entry
        self._collect_strings(self.root)
        self.string_index = {}
        objs = bytearray()
        for s in self.strings:
            if s in self.string_index:
                continue
            self.string_index[s] = len(self.string_index) + 1
            objs += struct.pack('<HH4xQ', self.string_index[s], 0, len(s)) + s + bytes(_pad8(len(s)) - len(s))
        size = max(_GCOL_MIN, _pad8(16 + len(objs) + 16))
        self.gcol = out.alloc(size)
        free = size - 16 - len(objs)
        out.put(self.gcol, b'GCOL' + struct.pack('<B3xQ', 1, size) + bytes(objs) + struct.pack('<HH4xQ', 0, 0, free))
        header, tree, heap = self._group(self.root)
        sb = b'\x89HDF\r\n\x1a\n' + bytes([0, 0, 0, 0, 0, 8, 8, 0])
        sb += struct.pack('<HHI', _LEAF_K, _NODE_K, 0)
        sb += struct.pack('<QQQQ', 0, UNDEF, len(out.buf), UNDEF)
        sb += struct.pack('<QQII', 0, header, 1, 0) + struct.pack('<QQ', tree, heap)
        out.put(0, sb)
        with open(path, 'wb') as fh:
            fh.write(out.buf)


def write_file(path, root):
    """Write the Group tree `root` as an HDF5 file."""
    _Writer(root).write(path)
