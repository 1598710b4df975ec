"""Checkpoint files in the reference's format without libhdf5.

The reference stores its state in HDF5 files written through the HDF5 C
library: ``<simname>_checkpoint_<k>.h5`` with ``vorticity/complex/<it>``
(compound {r, i} of rnumber, ``[ny][nz][nx/2+1][3]``),
``tracers0/state/<it>`` (``[np][3]`` f64) and ``tracers0/rhs/<it>``
(``[nsteps][np][3]`` f64) (bfps/cpp/field.cpp:164-356,
bfps/cpp/vorticity_equation.cpp:49-108,
bfps/cpp/particles/particles_output_hdf5.hpp:110-300).  libhdf5 and h5py
are not part of this build, so this module reads and writes the subset of
the HDF5 file format those files use:

* superblock version 0, 8-byte offsets and lengths, group leaf K 4 and
  internal K 16 (the library defaults);
* groups as symbol tables: a version-1 B-tree of symbol-table nodes plus a
  local heap of names;
* version-1 object headers; dataspace (v1, v2), datatype (fixed point,
  floating point, fixed-length strings, and the two-member compound {r, i}
  of floats that holds complex numbers), fill value, data layout v3
  (contiguous or compact) and modification-time messages; header
  continuations when reading.  Strings are written fixed-length (a Python
  str as ASCII bytes).  Reading also takes what h5py writes into the
  reference launcher's run file: variable-length strings (global heap
  collections) and chunked, unfiltered datasets (data layout v3 class 2,
  v1 B-tree of chunks), such as its extendable statistics series.

The writer lays files out as the C library does for these objects (the
same message set, flags and encodings, dataspaces with maximum dimensions,
early space allocation).  A fresh file is written in one pass.  An existing
file opened in mode "a" is appended to without moving anything in it: new
datasets and groups go to the end of the file; a dataset assigned again
with its type and shape is overwritten in place; each group whose entries
changed gets a new heap, symbol-table nodes and B-tree at the end, and its
symbol-table message (and the cached copy in its parent's entry) is patched
in place.  Anything outside the subset raises ``H5FormatError`` so that a
misread cannot pass silently.
"""
import os
import struct
import time

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K = 4        # symbol-table node holds up to 2 LEAF_K entries
INTERNAL_K = 16   # B-tree node holds up to 2 INTERNAL_K children


class H5FormatError(Exception):
    pass


# variable-length strings: read as Python bytes (an object array), stored as
# (length, global heap collection address, object index) = 16 bytes
VLEN_STR = np.dtype(object)
VLEN_BYTES = 16


# ---------------------------------------------------------------- datatypes
def _float_props(size):
    """(bit field bytes, properties) of an IEEE little-endian float."""
    if size == 4:
        return bytes([0x20, 31, 0]), struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
    if size == 8:
        return bytes([0x20, 63, 0]), struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
    raise H5FormatError("float size %d" % size)


def _encode_datatype(dt):
    """Datatype message body for a numpy dtype."""
    dt = np.dtype(dt)
    if dt.kind == "f":
        bits, props = _float_props(dt.itemsize)
        return bytes([0x11]) + bits + struct.pack("<I", dt.itemsize) + props
    if dt.kind in "iu":
        signed = 0x08 if dt.kind == "i" else 0x00
        return (bytes([0x10, signed, 0, 0]) + struct.pack("<I", dt.itemsize) +
                struct.pack("<HH", 0, 8 * dt.itemsize))
    if dt.kind == "S":
        # fixed-length ASCII string, null padded (numpy bytes_, as h5py
        # stores a Python bytes value)
        return bytes([0x13, 0x01, 0, 0]) + struct.pack("<I", dt.itemsize)
    if dt.kind == "c":
        half = dt.itemsize // 2
        member = _encode_datatype(np.dtype("<f%d" % half))
        body = b""
        for i, name in enumerate((b"r", b"i")):
            nm = name + b"\x00" * (8 - len(name))
            body += nm + struct.pack("<I", i * half) + bytes([0, 0, 0, 0]) + struct.pack("<I", 0) + b"\x00" * 4
            body += b"\x00" * 16 + member
        return bytes([0x16, 2, 0, 0]) + struct.pack("<I", dt.itemsize) + body
    raise H5FormatError("dtype %s" % dt)


def _decode_datatype(b, off):
    """numpy dtype of the datatype encoded at b[off:], and its length."""
    cls, ver = b[off] & 0x0F, b[off] >> 4
    bits = b[off + 1:off + 4]
    size = struct.unpack_from("<I", b, off + 4)[0]
    if cls == 0:  # fixed point
        if bits[0] & 0x01:
            raise H5FormatError("big-endian integers")
        return np.dtype("<%s%d" % ("i" if bits[0] & 0x08 else "u", size)), 8 + 4
    if cls == 1:  # floating point
        if bits[0] & 0x01 or bits[0] & 0x40:
            raise H5FormatError("non little-endian float")
        if size not in (4, 8):
            raise H5FormatError("float size %d" % size)
        return np.dtype("<f%d" % size), 8 + 12
    if cls == 3:  # fixed-length string (any padding, ASCII or UTF-8)
        return np.dtype("S%d" % size), 8
    if cls == 9:  # variable-length: strings only (h5py's str), read from the global heap
        if bits[0] & 0x0F != 1:
            raise H5FormatError("variable-length sequences")
        blen = _decode_datatype(b, off + 8)[1]
        return VLEN_STR, 8 + blen
    if cls == 6:  # compound: {r, i} floats -> complex
        nmemb = bits[0] | (bits[1] << 8)
        p = off + 8
        members = []
        for _ in range(nmemb):
            e = b.index(b"\x00", p)
            name = b[p:e].decode()
            if ver == 3:
                p = e + 1
                nb = 1
                while size >= (1 << (8 * nb)):
                    nb += 1
                moff = int.from_bytes(b[p:p + nb], "little")
                p += nb
            else:
                p += ((e - p) // 8 + 1) * 8
                moff = struct.unpack_from("<I", b, p)[0]
                p += 4
                if ver == 1:
                    if b[p] != 0:
                        raise H5FormatError("array members")
                    p += 28
            mdt, mlen = _decode_datatype(b, p)
            p += mlen
            members.append((name, moff, mdt))
        if (len(members) == 2 and [m[0] for m in members] == ["r", "i"] and members[0][2] == members[1][2]
                and members[0][2].kind == "f" and members[0][1] == 0 and members[1][1] == members[0][2].itemsize
                and size == 2 * members[0][2].itemsize):
            return np.dtype("<c%d" % size), p - off
        raise H5FormatError("compound type other than {r, i}: %s" % members)
    raise H5FormatError("datatype class %d" % cls)


# ---------------------------------------------------------------- reading
class _Group(dict):
    """A group read from a file: its entries, plus where its symbol-table
    message (stab_at) and the cached copy of its (B-tree, heap) addresses in
    its parent's entry (scratch_at; the superblock's for the root) are."""
    hdr = btree = heap = stab_at = scratch_at = None


class _Dataset:
    def __init__(self, dtype, shape, address=None, data=None, compact=None, chunks=None):
        self.dtype, self.shape = np.dtype(dtype), tuple(int(s) for s in shape)
        self.address, self.data, self.compact = address, data, compact
        self.chunks = chunks  # (B-tree address, chunk shape) of chunked storage


def _u(b, off, n):
    return int.from_bytes(b[off:off + n], "little")


class _Parser:
    def __init__(self, path):
        self.f = open(path, "rb")
        head = self.f.read(96)
        if head[:8] != SIGNATURE:
            raise H5FormatError("%s is not an HDF5 file" % path)
        if head[8] != 0:
            raise H5FormatError("superblock version %d" % head[8])
        if head[13] != 8 or head[14] != 8:
            raise H5FormatError("offset/length sizes %d/%d" % (head[13], head[14]))
        self.root = _u(head, 56 + 8, 8)
        self.eof = _u(head, 40, 8)

    def read(self, addr, n):
        self.f.seek(addr)
        b = self.f.read(n)
        if len(b) != n:
            raise H5FormatError("short read at %d" % addr)
        return b

    def messages(self, addr):
        h = self.read(addr, 16)
        if h[0] != 1:
            raise H5FormatError("object header version %d" % h[0])
        nmsg, hsize = _u(h, 2, 2), _u(h, 8, 4)
        blocks, out = [(addr + 16, hsize)], []
        while blocks:
            start, length = blocks.pop(0)
            b = self.read(start, length)
            p = 0
            while p + 8 <= length and len(out) < nmsg:
                mtype, msize, flags = _u(b, p, 2), _u(b, p + 2, 2), b[p + 4]
                body = b[p + 8:p + 8 + msize]
                if mtype == 0x10:
                    blocks.append((_u(body, 0, 8), _u(body, 8, 8)))
                if flags & 0x02 and mtype not in (0x00, 0x10):
                    raise H5FormatError("shared message type 0x%x" % mtype)
                out.append((mtype, body, start + p + 8))
                p += 8 + msize
        return out

    def object(self, addr, scratch_at=None):
        msgs = self.messages(addr)
        for t, body, at in msgs:
            if t == 0x11:
                g = self.group(_u(body, 0, 8), _u(body, 8, 8))
                g.hdr, g.stab_at, g.scratch_at = addr, at, scratch_at
                return g
        dtype = shape = layout = None
        for t, body, _ in msgs:
            if t == 0x01:
                ver, rank, flags = body[0], body[1], body[2]
                p = 8 if ver == 1 else 4
                if ver not in (1, 2):
                    raise H5FormatError("dataspace version %d" % ver)
                if ver == 2 and body[3] == 0:
                    rank = 0
                shape = tuple(_u(body, p + 8 * i, 8) for i in range(rank))
            elif t == 0x03:
                dtype = _decode_datatype(body, 0)[0]
            elif t == 0x08:
                if body[0] != 3:
                    raise H5FormatError("layout version %d" % body[0])
                if body[1] == 1:
                    layout = ("contiguous", _u(body, 2, 8), _u(body, 10, 8))
                elif body[1] == 0:
                    n = _u(body, 2, 2)
                    layout = ("compact", bytes(body[4:4 + n]))
                elif body[1] == 2:
                    # chunked: dimensionality (rank + 1: the last "dimension"
                    # is the element size), v1 B-tree address, chunk sizes
                    nd = body[2]
                    dims = tuple(_u(body, 11 + 4 * i, 4) for i in range(nd))
                    layout = ("chunked", _u(body, 3, 8), dims[:-1])
                else:
                    raise H5FormatError("layout class %d" % body[1])
            elif t == 0x0B:
                if body[1] != 0:  # number of filters (compression is outside the subset)
                    raise H5FormatError("filtered (compressed) dataset")
            elif t == 0x0C or t > 0x18:
                raise H5FormatError("object header message 0x%x" % t)
        if dtype is None or shape is None or layout is None:
            raise H5FormatError("object at %d is neither a group nor a dataset" % addr)
        if layout[0] == "compact":
            ds = _Dataset(dtype, shape, compact=layout[1])
        elif layout[0] == "chunked":
            ds = _Dataset(dtype, shape, chunks=(layout[1], layout[2]))
        else:
            ds = _Dataset(dtype, shape, address=None if layout[1] == UNDEF else layout[1])
        ds.hdr = addr
        return ds

    def group(self, btree, heap):
        h = self.read(heap, 32)
        if h[:4] != b"HEAP":
            raise H5FormatError("local heap signature")
        names = self.read(_u(h, 24, 8), _u(h, 8, 8))
        entries = {}
        self._btree(btree, names, entries)
        g = _Group()
        g.btree, g.heap = btree, heap
        for name, (addr, at) in entries.items():
            g[name] = self.object(addr, at)
        return g

    def _btree(self, addr, names, out):
        h = self.read(addr, 24)
        if h[:4] != b"TREE" or h[4] != 0:
            raise H5FormatError("group B-tree node")
        level, used = h[5], _u(h, 6, 2)
        body = self.read(addr + 24, 8 * (2 * used + 1))
        for i in range(used):
            child = _u(body, 8 + 16 * i, 8)
            if level > 0:
                self._btree(child, names, out)
            else:
                self._snod(child, names, out)

    def _snod(self, addr, names, out):
        h = self.read(addr, 8)
        if h[:4] != b"SNOD":
            raise H5FormatError("symbol table node signature")
        n = _u(h, 6, 2)
        body = self.read(addr + 8, 40 * n)
        for e in range(n):
            noff, obj = _u(body, 40 * e, 8), _u(body, 40 * e + 8, 8)
            end = names.index(b"\x00", noff)
            # the entry's cached (B-tree, heap) pair sits 24 bytes in
            out[names[noff:end].decode()] = (obj, addr + 8 + 40 * e + 24)

    def data(self, ds):
        vlen = ds.dtype == VLEN_STR
        est = VLEN_BYTES if vlen else ds.dtype.itemsize
        sdt = np.dtype("V%d" % VLEN_BYTES) if vlen else ds.dtype  # storage elements
        n = int(np.prod(ds.shape, dtype=np.int64)) * est
        if ds.chunks is not None:
            out = np.zeros(ds.shape, sdt)
            if ds.chunks[0] != UNDEF:
                self._chunks(ds.chunks[0], ds.chunks[1], est, sdt, out)
        elif ds.compact is not None:
            out = np.frombuffer(ds.compact[:n], sdt).reshape(ds.shape).copy()
        elif ds.address is None:
            out = np.zeros(ds.shape, sdt)
        else:
            out = np.frombuffer(self.read(ds.address, n), sdt).reshape(ds.shape).copy()
        return self._vlen_strings(out) if vlen else out

    def _chunks(self, addr, cshape, est, sdt, out):
        """Copy the chunks under v1 B-tree node addr (node type 1: raw data
        chunks; keys: chunk size, filter mask, chunk offsets) into out."""
        rank = len(cshape)
        h = self.read(addr, 24)
        if h[:4] != b"TREE" or h[4] != 1:
            raise H5FormatError("chunk B-tree node")
        level, used = h[5], _u(h, 6, 2)
        ksize = 8 + 8 * (rank + 1)
        body = self.read(addr + 24, used * (ksize + 8) + ksize)
        nbytes = int(np.prod(cshape, dtype=np.int64)) * est
        for i in range(used):
            k = i * (ksize + 8)
            child = _u(body, k + ksize, 8)
            if level > 0:
                self._chunks(child, cshape, est, sdt, out)
                continue
            if _u(body, k, 4) != nbytes or _u(body, k + 4, 4) != 0:
                raise H5FormatError("filtered or partial chunk")
            off = [_u(body, k + 8 + 8 * d, 8) for d in range(rank)]
            chunk = np.frombuffer(self.read(child, nbytes), sdt).reshape(cshape)
            dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, cshape, out.shape))
            src = tuple(slice(0, x.stop - x.start) for x in dst)
            out[dst] = chunk[src]

    def _vlen_strings(self, raw):
        """Variable-length strings: each element (length, global heap
        collection address, object index) names an object of a collection."""
        b = raw.tobytes()
        out = np.empty(raw.shape, dtype=object)
        flat = out.reshape(-1)
        cache = getattr(self, "_gcol", None)
        if cache is None:
            cache = self._gcol = {}
        for i in range(flat.size):
            length, coll, idx = _u(b, 16 * i, 4), _u(b, 16 * i + 4, 8), _u(b, 16 * i + 12, 4)
            if coll == 0 or coll == UNDEF:
                flat[i] = b""
                continue
            if coll not in cache:
                cache[coll] = self._collection(coll)
            flat[i] = cache[coll][idx][:length]
        return out[()] if out.ndim == 0 else out

    def _collection(self, addr):
        h = self.read(addr, 16)
        if h[:4] != b"GCOL":
            raise H5FormatError("global heap collection signature")
        size = _u(h, 8, 8)
        b = self.read(addr, size)
        objs, p = {}, 16
        while p + 16 <= size:
            idx, osize = _u(b, p, 2), _u(b, p + 8, 8)
            if idx == 0:
                break
            objs[idx] = bytes(b[p + 16:p + 16 + osize])
            p += 16 + _pad8(osize)
        return objs


# ---------------------------------------------------------------- writing
def _pad8(n):
    return (n + 7) // 8 * 8


def _msg(mtype, body, flags=0):
    body = body + b"\x00" * (_pad8(len(body)) - len(body))
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _object_header(messages):
    data = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(data)) + data


def _dataset_messages(shape, dtype, data_addr, nbytes, mtime):
    rank = len(shape)
    space = struct.pack("<BBB5x", 1, rank, 1) + b"".join(struct.pack("<Q", s) for s in shape) * 2
    return [
        _msg(0x01, space),
        _msg(0x03, _encode_datatype(dtype), flags=1),
        _msg(0x05, bytes([2, 1, 2, 1]) + struct.pack("<I", 0), flags=1),
        _msg(0x08, bytes([3, 1]) + struct.pack("<QQ", data_addr, nbytes), flags=1),
        _msg(0x12, struct.pack("<B3xI", 1, int(mtime))),
    ]


class _Writer:
    """Lays out a tree of groups ({name: subtree}) and arrays in one file."""

    def __init__(self, fh):
        self.fh = fh
        self.pos = 96
        self.mtime = time.time()

    def alloc(self, n):
        a = self.pos
        self.pos += _pad8(n)
        return a

    def put(self, addr, b):
        self.fh.seek(addr)
        self.fh.write(b)

    def write_object(self, node):
        """Returns (object header address, cache type, scratch bytes)."""
        if isinstance(node, dict):
            btree, heap = self.write_group(node)
            return self._group_header(btree, heap), 1, struct.pack("<QQ", btree, heap)
        arr = np.array(node, order="C")  # (ascontiguousarray would make a 0-d array 1-d)
        if arr.dtype.byteorder == ">":
            arr = arr.astype(arr.dtype.newbyteorder("<"))
        nbytes = arr.nbytes
        probe = _object_header(_dataset_messages(arr.shape, arr.dtype, 0, nbytes, self.mtime))
        hdr = self.alloc(len(probe))
        daddr = self.alloc(nbytes) if nbytes else UNDEF
        self.put(hdr, _object_header(_dataset_messages(arr.shape, arr.dtype, daddr, nbytes, self.mtime)))
        if nbytes:
            self.put(daddr, arr.tobytes())
        return hdr, 0, b"\x00" * 16

    def _group_header(self, btree, heap):
        h = _object_header([_msg(0x11, struct.pack("<QQ", btree, heap))])
        a = self.alloc(len(h))
        self.put(a, h)
        return a

    def write_group(self, tree):
        names = sorted(tree, key=lambda s: s.encode())
        # local heap: "" at offset 0, then each name, 8-byte aligned
        seg, offs = b"\x00" * 8, {}
        for nm in names:
            enc = nm.encode()
            if b"\x00" in enc or not enc or "/" in nm:
                raise H5FormatError("bad link name %r" % nm)
            offs[nm] = len(seg)
            seg += enc + b"\x00" * (_pad8(len(enc) + 1) - len(enc))
        heap = self.alloc(32)
        dseg = self.alloc(len(seg))
        self.put(heap, b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(seg), UNDEF, dseg))
        self.put(dseg, seg)
        children = {nm: self.write_object(tree[nm]) for nm in names}
        if not names:  # an empty group: a B-tree node without entries
            a = self.alloc(24 + 8 * (4 * INTERNAL_K + 1))
            self.put(a, b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", 0, UNDEF, UNDEF) +
                     b"\x00" * (8 * (4 * INTERNAL_K + 1)))
            return a, heap
        # symbol-table nodes of up to 2 LEAF_K entries, in name order
        leaves = []
        for i in range(0, len(names), 2 * LEAF_K):
            chunk = names[i:i + 2 * LEAF_K]
            body = b"".join(struct.pack("<QQI4x", offs[nm], children[nm][0], children[nm][1]) + children[nm][2]
                            for nm in chunk)
            body += b"\x00" * (40 * 2 * LEAF_K - len(body))
            a = self.alloc(8 + len(body))
            self.put(a, b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(chunk)) + body)
            leaves.append((a, offs[chunk[-1]]))
        # B-tree levels of up to 2 INTERNAL_K children
        level, nodes = 0, leaves
        while True:
            parents = []
            for i in range(0, len(nodes), 2 * INTERNAL_K):
                chunk = nodes[i:i + 2 * INTERNAL_K]
                body = struct.pack("<Q", 0)
                for child, last in chunk:
                    body += struct.pack("<QQ", child, last)
                body += b"\x00" * (8 * (4 * INTERNAL_K + 1) - len(body))
                a = self.alloc(24 + len(body))
                self.put(a, b"TREE" + bytes([0, level]) + struct.pack("<HQQ", len(chunk), UNDEF, UNDEF) + body)
                parents.append((a, chunk[-1][1]))
            if len(parents) == 1:
                return parents[0][0], heap
            nodes, level = parents, level + 1


def _write_file(path, tree):
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "wb") as fh:
        w = _Writer(fh)
        btree, heap = w.write_group(tree)
        root = w._group_header(btree, heap)
        eof = w.pos
        sb = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
        sb += struct.pack("<QQI4x", 0, root, 1) + struct.pack("<QQ", btree, heap)
        w.put(0, sb)
        fh.truncate(eof)
    os.replace(tmp, path)


class _Appender(_Writer):
    """Adds to an existing file without moving anything in it: new datasets
    (data, then object header) and new groups go to the end of the file; a
    dataset assigned again with its type and shape is overwritten in place;
    every group whose entries changed gets a new local heap, symbol-table
    nodes and B-tree at the end, its symbol-table message and the cached
    copy of it in its parent's entry (the superblock's for the root) being
    patched in place.  Superseded group structures stay behind as unused
    bytes."""

    def __init__(self, fh, eof):
        _Writer.__init__(self, fh)
        self.pos = _pad8(eof)

    def dataset(self, arr):
        arr = np.array(arr, order="C")
        if arr.dtype.byteorder == ">":
            arr = arr.astype(arr.dtype.newbyteorder("<"))
        daddr = self.alloc(arr.nbytes) if arr.nbytes else UNDEF
        if arr.nbytes:
            self.put(daddr, arr.tobytes())
        h = _object_header(_dataset_messages(arr.shape, arr.dtype, daddr, arr.nbytes, self.mtime))
        a = self.alloc(len(h))
        self.put(a, h)
        return _Dataset(arr.dtype, arr.shape, address=daddr if arr.nbytes else None), a

    def structure(self, entries):
        """heap + symbol-table nodes + B-tree of a group whose children are
        already in the file: entries {name: (header, cache type, scratch)}."""
        names = sorted(entries, key=lambda s: s.encode())
        seg, offs = b"\x00" * 8, {}
        for nm in names:
            enc = nm.encode()
            offs[nm] = len(seg)
            seg += enc + b"\x00" * (_pad8(len(enc) + 1) - len(enc))
        heap = self.alloc(32)
        dseg = self.alloc(len(seg))
        self.put(heap, b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(seg), UNDEF, dseg))
        self.put(dseg, seg)
        if not names:
            a = self.alloc(24 + 8 * (4 * INTERNAL_K + 1))
            self.put(a, b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", 0, UNDEF, UNDEF) +
                     b"\x00" * (8 * (4 * INTERNAL_K + 1)))
            return a, heap, {}
        scratch_at, leaves = {}, []
        for i in range(0, len(names), 2 * LEAF_K):
            chunk = names[i:i + 2 * LEAF_K]
            body = b"".join(struct.pack("<QQI4x", offs[nm], entries[nm][0], entries[nm][1]) + entries[nm][2]
                            for nm in chunk)
            body += b"\x00" * (40 * 2 * LEAF_K - len(body))
            a = self.alloc(8 + len(body))
            self.put(a, b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(chunk)) + body)
            for k, nm in enumerate(chunk):
                scratch_at[nm] = a + 8 + 40 * k + 24
            leaves.append((a, offs[chunk[-1]]))
        level, nodes = 0, leaves
        while True:
            parents = []
            for i in range(0, len(nodes), 2 * INTERNAL_K):
                chunk = nodes[i:i + 2 * INTERNAL_K]
                body = struct.pack("<Q", 0)
                for child, last in chunk:
                    body += struct.pack("<QQ", child, last)
                body += b"\x00" * (8 * (4 * INTERNAL_K + 1) - len(body))
                a = self.alloc(24 + len(body))
                self.put(a, b"TREE" + bytes([0, level]) + struct.pack("<HQQ", len(chunk), UNDEF, UNDEF) + body)
                parents.append((a, chunk[-1][1]))
            if len(parents) == 1:
                return parents[0][0], heap, scratch_at
            nodes, level = parents, level + 1

    def flush(self, g, pending):
        """Writes what changed below group g; returns g's replacement when
        its structure was rewritten (a _Group), else None.  `pending` maps
        id(group) to the set of its entries assigned since it was read."""
        changed = not isinstance(g, _Group)
        assigned = pending.get(id(g), set())
        for name in list(g):
            node = g[name]
            if isinstance(node, dict):
                new = self.flush(node, pending)
                if new is not None:
                    if not isinstance(node, _Group):  # a new group: its object header
                        h = _object_header([_msg(0x11, struct.pack("<QQ", new.btree, new.heap))])
                        new.hdr = self.alloc(len(h))
                        self.put(new.hdr, h)
                        new.stab_at = new.hdr + 16 + 8
                        changed = True
                    else:
                        new.hdr, new.stab_at, new.scratch_at = node.hdr, node.stab_at, node.scratch_at
                        self.patch(new)
                    g[name] = new
            elif name in assigned:
                arr = np.array(node, order="C")
                old = self.original(g, name)
                if (isinstance(old, _Dataset) and old.address is not None and old.compact is None
                        and old.dtype == arr.dtype and old.shape == arr.shape):
                    self.put(old.address, arr.astype(arr.dtype.newbyteorder("<")).tobytes())
                    g[name] = old
                else:
                    ds, hdr = self.dataset(arr)
                    ds.hdr = hdr
                    g[name] = ds
                    changed = True
        if not changed:
            return None
        entries = {}
        for name, node in g.items():
            if isinstance(node, dict):
                entries[name] = (node.hdr, 1, struct.pack("<QQ", node.btree, node.heap))
            else:
                entries[name] = (node.hdr, 0, b"\x00" * 16)
        btree, heap, scratch_at = self.structure(entries)
        out = _Group(g)
        out.btree, out.heap = btree, heap
        out.hdr, out.stab_at, out.scratch_at = getattr(g, "hdr", None), getattr(g, "stab_at", None), \
            getattr(g, "scratch_at", None)
        for name, node in out.items():
            if isinstance(node, _Group):
                node.scratch_at = scratch_at[name]
        return out

    def original(self, g, name):
        return self.originals.get((id(g), name))

    def patch(self, g):
        b = struct.pack("<QQ", g.btree, g.heap)
        self.put(g.stab_at, b)
        if g.scratch_at is not None:
            self.put(g.scratch_at, b)


# ---------------------------------------------------------------- interface
class File:
    """A checkpoint file: ``File(path, "r" | "a" | "w")``.  Datasets are
    addressed by slash-separated paths (``f["vorticity/complex/0"]``);
    complex arrays are stored as the reference's {r, i} compound.  In modes
    "a" and "w" assignments are written when the file is closed."""

    def __init__(self, path, mode="r"):
        if mode not in ("r", "a", "w"):
            raise ValueError("mode must be r, a or w")
        self.path, self.mode = path, mode
        self._tree, self._parser, self._dirty = {}, None, mode == "w"
        # entries assigned since the file was read: names per parent group,
        # and the dataset each replaced (for overwriting in place)
        self._assigned, self._originals = {}, {}
        if mode == "r" or (mode == "a" and os.path.exists(path)):
            self._parser = _Parser(path)
            self._tree = self._parser.object(self._parser.root, 80)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _node(self, path, create=False):
        parts = [p for p in path.split("/") if p]
        node = self._tree
        for i, p in enumerate(parts):
            if not isinstance(node, dict):
                raise KeyError(path)
            if p not in node:
                if not create:
                    raise KeyError(path)
                node[p] = {} if i + 1 < len(parts) else None
            if i + 1 < len(parts):
                node = node[p]
        return node, parts[-1] if parts else None

    def __contains__(self, path):
        try:
            parent, last = self._node(path)
            return last is None or last in parent
        except KeyError:
            return False

    def keys(self, path="/"):
        parent, last = self._node(path)
        node = parent if last is None else parent[last]
        if not isinstance(node, dict):
            raise KeyError("%s is a dataset" % path)
        return sorted(node)

    def __getitem__(self, path):
        parent, last = self._node(path)
        node = parent[last]
        if isinstance(node, dict):
            raise KeyError("%s is a group" % path)
        if isinstance(node, _Dataset):
            return self._parser.data(node)
        return np.array(node)

    def create_group(self, path):
        if self.mode == "r":
            raise IOError("file opened read-only")
        node = self._tree
        for p in [p for p in path.split("/") if p]:
            node = node.setdefault(p, {})
            if not isinstance(node, dict):
                raise KeyError("%s is a dataset" % path)
        self._dirty = True

    def __setitem__(self, path, value):
        if self.mode == "r":
            raise IOError("file opened read-only")
        parent, last = self._node(path, create=True)
        if isinstance(value, str):
            value = value.encode("ascii")
        arr = np.asarray(value)
        if arr.dtype.kind == "b":
            arr = arr.astype(np.int64)
        if arr.dtype.kind == "S" and arr.dtype.itemsize == 0:
            arr = arr.astype("S1")
        if arr.dtype.kind not in "fiucS":
            raise H5FormatError("dtype %s" % arr.dtype)
        if isinstance(parent.get(last), dict):
            raise KeyError("%s is a group" % path)
        key = (id(parent), last)
        if key not in self._originals:
            self._originals[key] = parent.get(last)
        self._assigned.setdefault(id(parent), set()).add(last)
        parent[last] = arr
        self._dirty = True

    def close(self):
        if self.mode != "r" and self._dirty and self._parser is not None:
            # an existing file: append, rewriting only the changed groups
            eof = max(self._parser.eof, os.path.getsize(self.path))
            self._parser.f.close()
            self._parser = None
            with open(self.path, "r+b") as fh:
                ap = _Appender(fh, eof)
                ap.originals = self._originals
                root = ap.flush(self._tree, self._assigned)
                if root is not None:
                    ap.patch(root)
                    self._tree = root
                if ap.pos != _pad8(eof):
                    ap.put(40, struct.pack("<Q", ap.pos))
            self._dirty = False
            self._assigned, self._originals = {}, {}
            return
        if self.mode != "r" and self._dirty:
            def materialise(node):
                if isinstance(node, dict):
                    return {k: materialise(v) for k, v in node.items()}
                if isinstance(node, _Dataset):
                    return self._parser.data(node)
                return node
            tree = materialise(self._tree)
            if self._parser is not None:
                self._parser.f.close()
                self._parser = None
            _write_file(self.path, tree)
            self._dirty = False
        if self._parser is not None:
            self._parser.f.close()
            self._parser = None
