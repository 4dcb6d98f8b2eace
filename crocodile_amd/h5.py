"""A NumPy reader and writer for the subset of HDF5 that crocodile's files
use, for machines without h5py or the HDF5 library.

crocodile keeps visibilities and kernels in HDF5 files written by h5py
(scripts/generate_wkernels.py, arl/test_support.py:292-396) and by the HDF5
C library (examples/grid/hdf5.c:416-478).  Those are classic-format files:

* superblock version 0 or 1, version-1 object headers (with continuation
  blocks), groups as symbol tables (version-1 B-tree of group nodes, symbol
  table nodes, local heap);
* datasets of little-endian integers, IEEE floats, fixed-length strings or
  compounds, complex numbers being h5py's / hdf5.c's compound {r, i}
  (hdf5.c:17-24); layouts compact, contiguous or chunked (version-1 B-tree of
  chunks, no filters), missing chunks reading as the fill value;
* attributes, with variable-length strings read from the global heap.

``File(path)`` opens a file read-only: ``f['wkern/0.1/-300.0/kern']`` is a
:class:`Dataset` (``.shape``, ``.dtype``, ``[()]`` reads it), groups list
their members (``keys()``, iteration, ``in``) and both carry ``.attrs``.
``write(path, tree)`` writes a file of that same form from a nested dict of
arrays (groups as dicts, datasets contiguous).  Anything outside the subset
(new-style groups of HDF5 1.8's latest format, filters such as compression,
big-endian data, references) raises NotImplementedError naming it.  The
file-format structures follow the published HDF5 File Format Specification
(version 2.0, sections III.A-IV.A).
"""

import struct

import numpy

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b'\x89HDF\r\n\x1a\n'

# object header message types (File Format Specification IV.A.2)
MSG_DATASPACE, MSG_LINKINFO, MSG_DATATYPE, MSG_FILL_OLD, MSG_FILL = 0x1, 0x2, 0x3, 0x4, 0x5
MSG_LINK, MSG_LAYOUT, MSG_FILTERS, MSG_ATTRIBUTE = 0x6, 0x8, 0xB, 0xC
MSG_CONTINUATION, MSG_SYMBOL_TABLE = 0x10, 0x11


def _u(buf, off, n):
    return int.from_bytes(buf[off:off + n], 'little')


class _Reader:
    """Byte access to the file (memory-mapped); addresses are relative to
    the base address (the superblock's offset behind a user block)."""

    def __init__(self, path):
        self.mm = numpy.memmap(path, dtype=numpy.uint8, mode='r')
        self.base = 0

    def bytes(self, off, n):
        off += self.base
        if off < 0 or off + n > self.mm.shape[0]:
            raise ValueError('HDF5 structure at %d+%d lies beyond the end of the file' % (off, n))
        return bytes(self.mm[off:off + n])


# ---------------------------------------------------------------------------
# datatypes and dataspaces
# ---------------------------------------------------------------------------

def _parse_datatype(b, off=0):
    """(numpy dtype, encoded length, info) of the datatype message at b[off]."""
    cv = b[off]
    cls, ver = cv & 0xF, cv >> 4
    bits = _u(b, off + 1, 3)
    size = _u(b, off + 4, 4)
    p = off + 8
    if cls == 0:                                  # fixed-point
        if bits & 1:
            raise NotImplementedError('big-endian integers')
        signed = bool(bits & 8)
        return numpy.dtype('<%s%d' % ('i' if signed else 'u', size)), 8 + 4, None
    if cls == 1:                                  # floating point
        if bits & 1 or bits & 0x40:
            raise NotImplementedError('big-endian or VAX floats')
        return numpy.dtype('<f%d' % size), 8 + 12, None
    if cls == 3:                                  # fixed-length string
        return numpy.dtype('S%d' % size), 8, None
    if cls == 6:                                  # compound
        n = bits & 0xFFFF
        names, formats, offsets = [], [], []
        for _ in range(n):
            end = b.index(b'\0', p)
            name = b[p:end].decode('utf-8')
            if ver < 3:
                p += ((end - p) // 8 + 1) * 8     # name padded to a multiple of 8
                moff = _u(b, p, 4)
                p += 4
                if ver == 1:
                    p += 1 + 3 + 4 + 4 + 16       # dimensionality, permutation, sizes
            else:
                p = end + 1
                nb = 1 if size < 256 else 2 if size < 65536 else 4 if size < 1 << 32 else 8
                moff = _u(b, p, nb)
                p += nb
            mdt, mlen, _ = _parse_datatype(b, p)
            p += mlen
            names.append(name)
            formats.append(mdt)
            offsets.append(moff)
        dt = numpy.dtype({'names': names, 'formats': formats, 'offsets': offsets,
                          'itemsize': size})
        return dt, p - off, None
    if cls == 9:                                  # variable length
        base, blen, _ = _parse_datatype(b, p)
        kind = bits & 0xF
        return numpy.dtype(object), 8 + blen, ('vlen_str' if kind == 1 else 'vlen_seq', base)
    if cls == 10:                                 # array
        raise NotImplementedError('array datatypes')
    raise NotImplementedError('HDF5 datatype class %d' % cls)


def _complex_view(a):
    """h5py's complex convention: a compound of two equal floats named r, i."""
    dt = a.dtype
    if dt.names == ('r', 'i') and dt.fields['r'][0] == dt.fields['i'][0] and \
            dt.fields['r'][0].kind == 'f' and dt.fields['r'][1] == 0 and \
            dt.fields['i'][1] == dt.fields['r'][0].itemsize and \
            dt.itemsize == 2 * dt.fields['r'][0].itemsize:
        ct = {4: numpy.complex64, 8: numpy.complex128}[dt.fields['r'][0].itemsize]
        return numpy.ascontiguousarray(a).view(ct).reshape(a.shape)
    return a


def _parse_dataspace(b, off=0):
    ver, rank, flags = b[off], b[off + 1], b[off + 2]
    if ver == 1:
        p = off + 8
        dims = [_u(b, p + 8 * k, 8) for k in range(rank)]
    elif ver == 2:
        if b[off + 3] == 2:                       # null dataspace
            return None
        p = off + 4
        dims = [_u(b, p + 8 * k, 8) for k in range(rank)]
    else:
        raise NotImplementedError('dataspace message version %d' % ver)
    return tuple(dims)


# ---------------------------------------------------------------------------
# objects
# ---------------------------------------------------------------------------

class _Object:
    def __init__(self, f, addr, name):
        self._f, self._addr, self.name = f, addr, name
        self.msgs = f._messages(addr)
        self._attrs = None

    @property
    def attrs(self):
        if self._attrs is None:
            self._attrs = {}
            for t, b in self.msgs:
                if t == MSG_ATTRIBUTE:
                    k, v = self._f._attribute(b)
                    self._attrs[k] = v
        return self._attrs


class Dataset(_Object):
    """A dataset: shape, dtype, and its data by ``ds[()]`` (``read()``)."""

    def __init__(self, f, addr, name):
        super().__init__(f, addr, name)
        self._space = self._dtype = self._layout = self._fill = None
        for t, b in self.msgs:
            if t == MSG_DATASPACE:
                self._space = _parse_dataspace(b)
            elif t == MSG_DATATYPE:
                self._dtype = _parse_datatype(b)
            elif t == MSG_LAYOUT:
                self._layout = b
            elif t == MSG_FILL and len(b) >= 2:
                self._fill = b
            elif t == MSG_FILTERS:
                raise NotImplementedError('filtered (compressed) dataset %s' % name)
        if self._dtype is None or self._layout is None:
            raise ValueError('%s is not a dataset' % name)

    @property
    def shape(self):
        return () if self._space is None else self._space

    @property
    def dtype(self):
        dt = self._dtype[0]
        if dt.names == ('r', 'i'):
            return _complex_view(numpy.zeros(1, dt)).dtype
        return dt

    def __len__(self):
        return self.shape[0]

    def __getitem__(self, key):
        a = self.read()
        return a if key == () or key is Ellipsis else a[key]

    def __array__(self, dtype=None, copy=None):
        a = self.read()
        return a if dtype is None else a.astype(dtype)

    def _fill_value(self, dt):
        out = numpy.zeros((), dt)
        b = self._fill
        if b is not None and b[0] in (2, 3):
            if b[0] == 2 and b[3] == 1 and len(b) >= 8:
                n = _u(b, 4, 4)
                if n == dt.itemsize:
                    out = numpy.frombuffer(b[8:8 + n], dt)[0]
            elif b[0] == 3:
                fl = b[1]
                if fl & 0x20 and len(b) >= 6:
                    n = _u(b, 2, 4)
                    if n == dt.itemsize:
                        out = numpy.frombuffer(b[6:6 + n], dt)[0]
        return out

    def read(self):
        dt, _, vlen = self._dtype
        if vlen is not None:
            raise NotImplementedError('variable-length dataset %s' % self.name)
        shape = self.shape
        count = int(numpy.prod(shape)) if shape else 1
        b = self._layout
        ver = b[0]
        f = self._f
        if ver == 3:
            cls = b[1]
            if cls == 0:
                compact = b[4:4 + _u(b, 2, 2)]
            elif cls == 1:
                addr = _u(b, 2, 8)
            elif cls == 2:
                nd, btree = b[2], _u(b, 3, 8)
                cdims = [_u(b, 11 + 4 * k, 4) for k in range(nd)]
        elif ver in (1, 2):
            # version 1 / 2: dimensionality, class, 5 reserved bytes, the
            # address (not for compact), the dimension sizes (4 bytes each),
            # then the element size (chunked) or the compact data
            nd, cls = b[1], b[2]
            p = 8 if cls == 0 else 16
            addr = btree = _u(b, 8, 8)
            cdims = [_u(b, p + 4 * k, 4) for k in range(nd)]
            p += 4 * nd
            if cls == 0:
                compact = b[p + 4:p + 4 + _u(b, p, 4)]
            elif cls == 2:
                cdims = cdims[:len(shape)] + [_u(b, p, 4)]
                nd = len(cdims)
        else:
            raise NotImplementedError('data layout message version %d in %s' % (ver, self.name))
        if cls == 0:                                  # compact
            a = numpy.frombuffer(compact, dt, count)
        elif cls == 1:                                # contiguous
            if addr == UNDEF:
                a = numpy.full(count, self._fill_value(dt), dt)
            else:
                a = numpy.frombuffer(f._r.bytes(addr, count * dt.itemsize), dt, count)
        elif cls == 2:                                # chunked, version-1 B-tree index
            a = numpy.full(shape, self._fill_value(dt), dt)
            if btree != UNDEF:
                for coff, caddr, csize in f._chunks(btree, nd):
                    cshape = tuple(cdims[:-1])
                    chunk = numpy.frombuffer(f._r.bytes(caddr, int(numpy.prod(cshape)) *
                                                        dt.itemsize), dt).reshape(cshape)
                    sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(coff[:-1], cshape,
                                                                            shape))
                    a[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
            a = a.reshape(-1)
        else:
            raise NotImplementedError('layout class %d in %s' % (cls, self.name))
        a = numpy.array(a).reshape(shape)
        return _complex_view(a)


class Group(_Object):
    """A group: members by name (nested paths allowed), keys(), items()."""

    def __init__(self, f, addr, name):
        super().__init__(f, addr, name)
        self._members = None

    def _links(self):
        if self._members is None:
            self._members = {}
            st = [b for t, b in self.msgs if t == MSG_SYMBOL_TABLE]
            if st:
                btree, heap = _u(st[0], 0, 8), _u(st[0], 8, 8)
                self._members = self._f._symbol_table(btree, heap)
            elif any(t in (MSG_LINK, MSG_LINKINFO) for t, _ in self.msgs):
                raise NotImplementedError('new-style (link message) group %s' % self.name)
        return self._members

    def keys(self):
        return list(self._links().keys())

    def __iter__(self):
        return iter(self.keys())

    def __len__(self):
        return len(self._links())

    def __contains__(self, path):
        try:
            self[path]
            return True
        except KeyError:
            return False

    def __getitem__(self, path):
        obj = self
        parts = [p for p in path.split('/') if p]
        if not parts:
            return self
        for k, part in enumerate(parts):
            if not isinstance(obj, Group):
                raise KeyError(path)
            links = obj._links()
            if part not in links:
                raise KeyError(path)
            name = obj.name.rstrip('/') + '/' + part
            obj = obj._f._open(links[part], name)
        return obj

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def values(self):
        return [self[k] for k in self.keys()]


class File(Group):
    """A classic-format HDF5 file, read-only."""

    def __init__(self, path):
        self._r = _Reader(path)
        # the superblock sits at 0 or, behind a user block, at 512, 1024, ...
        at = 0
        while at + 96 <= self._r.mm.shape[0] and self._r.bytes(at, 8) != SIGNATURE:
            at = 512 if at == 0 else 2 * at
        if at + 96 > self._r.mm.shape[0]:
            raise ValueError('%s is not an HDF5 file (no superblock signature)' % path)
        sb = self._r.bytes(at, 96)
        ver = sb[8]
        if ver not in (0, 1):
            raise NotImplementedError('superblock version %d (HDF5 1.8 "latest" format)' % ver)
        if sb[13] != 8 or sb[14] != 8:
            raise NotImplementedError('offsets / lengths of %d / %d bytes' % (sb[13], sb[14]))
        p = 24 + (4 if ver == 1 else 0)
        self.base = _u(sb, p, 8)
        self._r.base = self.base if self.base != UNDEF else at
        root_entry = p + 32
        root = _u(sb, root_entry + 8, 8)
        self._cache = {}
        super().__init__(self, root, '/')

    def _open(self, addr, name):
        if addr in self._cache:
            return self._cache[addr]
        msgs = self._messages(addr)
        types = {t for t, _ in msgs}
        if MSG_SYMBOL_TABLE in types or MSG_LINKINFO in types or MSG_LINK in types:
            obj = Group(self, addr, name)
        elif MSG_LAYOUT in types:
            obj = Dataset(self, addr, name)
        else:
            raise NotImplementedError('object %s is neither a group nor a dataset' % name)
        self._cache[addr] = obj
        return obj

    def _messages(self, addr):
        """[(type, data)] of a version-1 object header, continuations followed."""
        h = self._r.bytes(addr, 16)
        if h[0] != 1:
            if h[:4] == b'OHDR':
                raise NotImplementedError('version-2 object header at %d' % addr)
            raise ValueError('no object header at %d' % addr)
        nmsg, size = _u(h, 2, 2), _u(h, 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            b = self._r.bytes(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsg:
                t, n = _u(b, p, 2), _u(b, p + 2, 2)
                data = b[p + 8:p + 8 + n]
                if t == MSG_CONTINUATION:
                    blocks.append((_u(data, 0, 8), _u(data, 8, 8)))
                out.append((t, data))
                p += 8 + n
        return [(t, d) for t, d in out if t not in (0, MSG_CONTINUATION)]

    def _heap_string(self, heap, off):
        h = self._r.bytes(heap, 32)
        if h[:4] != b'HEAP':
            raise ValueError('no local heap at %d' % heap)
        data, size = _u(h, 24, 8), _u(h, 8, 8)
        if off >= size:
            raise ValueError('name offset %d beyond the local heap at %d' % (off, heap))
        b = self._r.bytes(data + off, size - off)
        return b[:b.index(b'\0')].decode('utf-8')

    def _symbol_table(self, btree, heap):
        out = {}

        def walk(node):
            h = self._r.bytes(node, 24)
            if h[:4] != b'TREE' or h[4] != 0:
                raise ValueError('no group B-tree node at %d' % node)
            level, used = h[5], _u(h, 6, 2)
            b = self._r.bytes(node + 24, 8 * (2 * used + 1))
            for k in range(used):
                child = _u(b, 8 + 16 * k, 8)
                if level > 0:
                    walk(child)
                else:
                    s = self._r.bytes(child, 8)
                    if s[:4] != b'SNOD':
                        raise ValueError('no symbol table node at %d' % child)
                    ns = _u(s, 6, 2)
                    e = self._r.bytes(child + 8, 40 * ns)
                    for i in range(ns):
                        name = self._heap_string(heap, _u(e, 40 * i, 8))
                        out[name] = _u(e, 40 * i + 8, 8)

        if btree != UNDEF:
            walk(btree)
        return dict(sorted(out.items()))

    def _chunks(self, node, nd):
        """(offsets, address, size) of every chunk under a chunk B-tree node."""
        h = self._r.bytes(node, 24)
        if h[:4] != b'TREE' or h[4] != 1:
            raise ValueError('no chunk B-tree node at %d' % node)
        level, used = h[5], _u(h, 6, 2)
        ksize = 8 + 8 * nd
        b = self._r.bytes(node + 24, used * (ksize + 8) + ksize)
        for k in range(used):
            key = k * (ksize + 8)
            csize, fmask = _u(b, key, 4), _u(b, key + 4, 4)
            offs = [_u(b, key + 8 + 8 * d, 8) for d in range(nd)]
            child = _u(b, key + ksize, 8)
            if level > 0:
                yield from self._chunks(child, nd)
            else:
                if fmask:
                    raise NotImplementedError('filtered chunk')
                yield offs, child, csize

    def _vlen_string(self, b, off):
        n, coll, idx = _u(b, off, 4), _u(b, off + 4, 8), _u(b, off + 12, 4)
        h = self._r.bytes(coll, 16)
        if h[:4] != b'GCOL':
            raise ValueError('no global heap collection at %d' % coll)
        size = _u(h, 8, 8)
        c = self._r.bytes(coll, size)
        p = 16
        while p + 16 <= size:
            i, osz = _u(c, p, 2), _u(c, p + 8, 8)
            if i == 0:
                break
            if i == idx:
                return c[p + 16:p + 16 + min(n, osz)].decode('utf-8')
            p += 16 + ((osz + 7) // 8) * 8
        raise ValueError('global heap object %d not found' % idx)

    def _attribute(self, b):
        ver = b[0]
        nsz, tsz, ssz = _u(b, 2, 2), _u(b, 4, 2), _u(b, 6, 2)
        if ver == 1:
            pad = lambda n: ((n + 7) // 8) * 8      # noqa: E731
            p = 8
            name = b[p:p + nsz].rstrip(b'\0').decode('utf-8')
            p += pad(nsz)
            dt, _, vlen = _parse_datatype(b, p)
            p += pad(tsz)
            shape = _parse_dataspace(b, p)
            p += pad(ssz)
        elif ver in (2, 3):
            p = 8 + (1 if ver == 3 else 0)
            name = b[p:p + nsz].rstrip(b'\0').decode('utf-8')
            p += nsz
            dt, _, vlen = _parse_datatype(b, p)
            p += tsz
            shape = _parse_dataspace(b, p)
            p += ssz
        else:
            raise NotImplementedError('attribute message version %d' % ver)
        shape = shape or ()
        count = int(numpy.prod(shape)) if shape else 1
        if vlen is not None:
            if vlen[0] != 'vlen_str':
                raise NotImplementedError('variable-length sequence attribute %s' % name)
            vals = [self._vlen_string(b, p + 16 * k) for k in range(count)]
            val = vals[0] if not shape else numpy.array(vals, dtype=object).reshape(shape)
            return name, val
        a = _complex_view(numpy.frombuffer(b[p:p + count * dt.itemsize], dt, count).copy())
        if dt.kind == 'S':
            a = numpy.char.decode(numpy.char.rstrip(a, b'\0'), 'utf-8')
        return name, (a[0] if not shape else a.reshape(shape))


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------

def _datatype_message(dt):
    """Encoded datatype message of a NumPy dtype (little-endian)."""
    dt = numpy.dtype(dt)
    if dt.kind == 'c':
        half = dt.itemsize // 2
        f = _datatype_message(numpy.dtype('<f%d' % half))
        mem = b''
        for i, name in enumerate((b'r', b'i')):
            nm = name + b'\0' * (8 - len(name))
            mem += nm + struct.pack('<I', i * half) + b'\0' * 28 + f
        return bytes([0x16]) + (2).to_bytes(3, 'little') + struct.pack('<I', dt.itemsize) + mem
    if dt.kind == 'f':
        if dt.itemsize == 8:
            props = struct.pack('<HHBBBBI', 0, 64, 52, 11, 0, 52, 1023)
        elif dt.itemsize == 4:
            props = struct.pack('<HHBBBBI', 0, 32, 23, 8, 0, 23, 127)
        else:
            raise NotImplementedError('float of %d bytes' % dt.itemsize)
        # bit field: little-endian, implied mantissa MSB (normalisation 2),
        # sign at bit size*8 - 1
        bits = (2 << 4) | ((dt.itemsize * 8 - 1) << 8)
        return bytes([0x11]) + bits.to_bytes(3, 'little') + struct.pack('<I', dt.itemsize) + props
    if dt.kind in 'iu':
        bits = 8 if dt.kind == 'i' else 0
        return (bytes([0x10]) + bits.to_bytes(3, 'little') + struct.pack('<I', dt.itemsize) +
                struct.pack('<HH', 0, dt.itemsize * 8))
    if dt.kind == 'S':
        return bytes([0x13]) + (0).to_bytes(3, 'little') + struct.pack('<I', dt.itemsize)
    raise NotImplementedError('dtype %s' % dt)


def _pad8(b):
    return b + b'\0' * (-len(b) % 8)


class VlenStr(str):
    """An attribute value to be written as a variable-length string (as h5py
    stores Python strings), its bytes in a global heap collection."""


class _Writer:
    def __init__(self):
        self.buf = bytearray(b'\0' * 96)        # superblock, filled in last

    def alloc(self, data):
        addr = len(self.buf)
        self.buf += _pad8(bytes(data))
        return addr

    def object_header(self, msgs):
        body = b''
        for t, d in msgs:
            d = _pad8(d)
            body += struct.pack('<HHB3x', t, len(d), 0) + d
        head = struct.pack('<BBHII', 1, 0, len(msgs), 1, len(body)) + b'\0' * 4
        return self.alloc(head + body)

    @staticmethod
    def _space(shape):
        return struct.pack('<BBB5x', 1, len(shape), 0) + b''.join(struct.pack('<Q', s)
                                                                for s in shape)

    def attributes(self, attrs):
        """Version-1 attribute messages: strings as fixed-length
        null-terminated strings, everything else as a NumPy array."""
        msgs = []
        for name, v in (attrs or {}).items():
            if isinstance(v, VlenStr):
                raw = v.encode('utf-8')
                # one-object global heap collection, the rest a free block
                obj = struct.pack('<HH4xQ', 1, 1, len(raw)) + _pad8(raw)
                size = max(4096, 16 + len(obj) + 16)
                free = struct.pack('<HH4xQ', 0, 0, size - 16 - len(obj))
                coll = b'GCOL' + struct.pack('<B3xQ', 1, size) + obj + free
                caddr = self.alloc(coll + b'\0' * (size - len(coll)))
                base = bytes([0x10]) + (0).to_bytes(3, 'little') + struct.pack('<IHH', 1, 0, 8)
                dt_msg = (bytes([0x19]) + (1 | (1 << 8)).to_bytes(3, 'little') +
                          struct.pack('<I', 16) + base)
                space, data = self._space(()), struct.pack('<IQI', len(raw), caddr, 1)
            elif isinstance(v, str):
                raw = v.encode('utf-8') + b'\0'
                dt_msg = bytes([0x13]) + (0).to_bytes(3, 'little') + struct.pack('<I', len(raw))
                space, data = self._space(()), raw
            else:
                a = numpy.asarray(v, order='C')
                dt_msg, space, data = _datatype_message(a.dtype), self._space(a.shape), a.tobytes()
            nm = name.encode('utf-8') + b'\0'
            body = (struct.pack('<BBHHH', 1, 0, len(nm), len(dt_msg), len(space)) +
                    _pad8(nm) + _pad8(dt_msg) + _pad8(space) + data)
            msgs.append((MSG_ATTRIBUTE, body))
        return msgs

    def chunked(self, a, chunks):
        """Layout message of ``a`` stored in chunks of shape ``chunks`` under
        a version-1 B-tree of chunk nodes (at most 64 entries each), edge
        chunks padded with zeros."""
        nd = a.ndim + 1
        grid = [range(0, max(n, 1), c) for n, c in zip(a.shape, chunks)]
        entries = []
        for off in numpy.ndindex(*[len(g) for g in grid]):
            o = [grid[d][off[d]] for d in range(a.ndim)]
            block = numpy.zeros(chunks, a.dtype)
            part = a[tuple(slice(x, x + c) for x, c in zip(o, chunks))]
            block[tuple(slice(0, n) for n in part.shape)] = part
            raw = block.tobytes()
            entries.append((o + [0], self.alloc(raw), len(raw)))

        def key(offs, size):
            return struct.pack('<II', size, 0) + b''.join(struct.pack('<Q', x) for x in offs)

        end = [n for n in a.shape] + [0]
        level, nodes = 0, [(e[0], e[1], e[2]) for e in entries]
        while True:
            new = []
            for i in range(0, len(nodes), 64):
                g = nodes[i:i + 64]
                body = b''
                for offs, child, size in g:
                    body += key(offs, size) + struct.pack('<Q', child)
                nxt = nodes[i + 64][0] if i + 64 < len(nodes) else end
                body += key(nxt, 0)
                addr = self.alloc(b'TREE' + struct.pack('<BBHQQ', 1, level, len(g), UNDEF,
                                                        UNDEF) + body)
                new.append((g[0][0], addr, 0))
            if len(new) == 1:
                root = new[0][1]
                break
            nodes, level = new, level + 1
        dims = b''.join(struct.pack('<I', c) for c in chunks) + struct.pack('<I', a.dtype.itemsize)
        return struct.pack('<BBBQ', 3, 2, nd, root) + dims

    def dataset(self, a, attrs=None, chunks=None):
        a = numpy.asarray(a, order='C')       # keeps a 0-d array 0-d
        if a.dtype.byteorder == '>':
            a = a.astype(a.dtype.newbyteorder('<'))
        if a.dtype == numpy.bool_:
            a = a.astype(numpy.uint8)
        space = struct.pack('<BBB5x', 1, a.ndim, 0) + b''.join(struct.pack('<Q', s)
                                                               for s in a.shape)
        if chunks is not None:
            if len(chunks) != a.ndim or min(chunks, default=1) < 1:
                raise ValueError('chunk shape %s does not fit the array' % (chunks,))
            layout = self.chunked(a, tuple(int(c) for c in chunks))
        else:
            addr = self.alloc(a.tobytes()) if a.nbytes else UNDEF
            layout = struct.pack('<BBQQ', 3, 1, addr, a.nbytes)
        fill = struct.pack('<BBBB', 2, 2, 1, 0)      # version 2, late allocation, never written, undefined
        return self.object_header([(MSG_DATASPACE, space), (MSG_DATATYPE, _datatype_message(a.dtype)),
                                   (MSG_FILL, fill), (MSG_LAYOUT, layout)] +
                                  self.attributes(attrs))

    def group(self, tree, path='', attrs=None, chunks=None):
        """Symbol-table group of the (name -> dict | array) tree; ``attrs``
        {object path: {name: value}}, ``chunks`` {dataset path: shape}."""
        attrs, chunks = attrs or {}, chunks or {}
        names = sorted(tree)
        # local heap: empty string at 0, then the names
        heap_data = b'\0' * 8
        offs = {}
        for n in names:
            offs[n] = len(heap_data)
            heap_data += _pad8(n.encode('utf-8') + b'\0')
        heap_data = _pad8(heap_data)
        children = {}
        for n in names:
            sub = path + '/' + n
            children[n] = (self.group(tree[n], sub, attrs, chunks) if isinstance(tree[n], dict)
                           else self.dataset(tree[n], attrs.get(sub), chunks.get(sub)))
        data_addr = self.alloc(heap_data)
        heap = self.alloc(b'HEAP' + struct.pack('<B3xQQQ', 0, len(heap_data), UNDEF, data_addr))
        # symbol table nodes of at most 2 K = 8 entries (group leaf node K 4)
        snods = []
        for i in range(0, max(len(names), 1), 8):
            part = names[i:i + 8]
            ent = b''.join(struct.pack('<QQII16x', offs[n], children[n], 0, 0) for n in part)
            ent += b'\0' * 40 * (8 - len(part))
            snods.append((part[-1] if part else '', self.alloc(b'SNOD' + struct.pack(
                '<BBH', 1, 0, len(part)) + ent)))
        if not names:
            snods = []
        # B-tree of group nodes: at most 2 K = 32 children per node (K 16)
        level, nodes = 0, [(offs[last] if last else 0, a) for last, a in snods]
        first_key = 0
        while True:
            groups = [nodes[i:i + 32] for i in range(0, max(len(nodes), 1), 32)] if nodes else [[]]
            new = []
            for g in groups:
                body = struct.pack('<Q', first_key)
                for key, child in g:
                    body += struct.pack('<QQ', child, key)
                body += b'\0' * 16 * (32 - len(g))
                addr = self.alloc(b'TREE' + struct.pack('<BBHQQ', 0, level, len(g), UNDEF, UNDEF) +
                                  body)
                new.append((g[-1][0] if g else 0, addr))
            if len(new) == 1:
                root = new[0][1]
                break
            nodes, level = new, level + 1
        stab = struct.pack('<QQ', root, heap)
        return self.object_header([(MSG_SYMBOL_TABLE, stab)] +
                                  self.attributes(attrs.get(path or '/')))

    def finish(self, root):
        # root entry without cached group addresses (cache type 0)
        root_entry = struct.pack('<QQII', 0, root, 0, 0) + b'\0' * 16
        sb = (SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack('<HHI', 4, 16, 0) +
              struct.pack('<QQQQ', 0, UNDEF, len(self.buf), UNDEF) + root_entry)
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def _tree(flat):
    """Nested dict of a {path: array} or already nested mapping."""
    out = {}
    for k, v in flat.items():
        if isinstance(v, dict):
            v = _tree(v)
        parts = [p for p in k.split('/') if p]
        d = out
        for p in parts[:-1]:
            d = d.setdefault(p, {})
        if parts[-1] in d and isinstance(d[parts[-1]], dict) and isinstance(v, dict):
            d[parts[-1]].update(v)
        else:
            d[parts[-1]] = v
    return out


def write(path, tree, attrs=None, chunks=None):
    """Write ``tree`` ({name or path: array or nested dict}) as a classic
    HDF5 file: groups as symbol tables, datasets contiguous unless ``chunks``
    {dataset path: chunk shape} says otherwise (as hdf5.c:449-466 stores
    ``vis``); complex arrays as the {r, i} compound.  ``attrs``: {object
    path ('/' for the root, '/vis', '/vis/uvw', ...): {name: value}}, a
    value a str (fixed-length string), :class:`VlenStr` (variable-length
    string, as h5py writes str) or an array."""
    def norm(d):
        return {('/' + k.strip('/')) if k.strip('/') else '/': v for k, v in (d or {}).items()}
    w = _Writer()
    root = w.group(_tree(tree), '', norm(attrs), norm(chunks))
    with open(path, 'wb') as fh:
        fh.write(w.finish(root))
