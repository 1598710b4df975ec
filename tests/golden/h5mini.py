"""Minimal read-only HDF5 reader (superblock v0, symbol-table groups,
object header v1, contiguous or unallocated datasets).

libhdf5 / h5py are not available in this environment, so the reference's
golden checkpoint (``bfps/test/B32p1e4_checkpoint_0.h5``) is decoded with
this reader.  Only the subset of the format that file uses is implemented:

* superblock version 0, 8-byte offsets and lengths;
* groups stored as v1 B-trees of symbol-table nodes plus a local heap;
* object header version 1 with dataspace (v1/v2), datatype (fixed-point,
  floating-point, compound of floats), and data-layout (v3 contiguous)
  messages, plus header-continuation messages.

Anything else raises ``NotImplementedError`` so that a silent misread is
impossible.
"""
import struct

import numpy as np


class H5Error(Exception):
    pass


class _Reader:
    def __init__(self, buf):
        self.b = buf

    def u(self, off, n):
        return int.from_bytes(self.b[off:off + n], "little")


class Dataset:
    def __init__(self, f, name, shape, dtype, address, size):
        self.file = f
        self.name = name
        self.shape = shape
        self.dtype = dtype
        self.address = address
        self.size = size

    def read(self):
        n = int(np.prod(self.shape)) if self.shape else 1
        nbytes = n * self.dtype.itemsize
        if self.address is None:
            # never written: HDF5 returns the fill value (0 by default)
            return np.zeros(self.shape, dtype=self.dtype)
        if self.size != nbytes:
            raise H5Error("size mismatch for %s" % self.name)
        raw = self.file.buf[self.address:self.address + nbytes]
        return np.frombuffer(raw, dtype=self.dtype).reshape(self.shape).copy()


class File:
    UNDEF = 0xFFFFFFFFFFFFFFFF

    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        r = _Reader(self.buf)
        if self.buf[:8] != b"\x89HDF\r\n\x1a\n":
            raise H5Error("not an HDF5 file")
        if self.buf[8] != 0:
            raise NotImplementedError("superblock version %d" % self.buf[8])
        if self.buf[13] != 8 or self.buf[14] != 8:
            raise NotImplementedError("offset/length size")
        # 8 sig + 8 version bytes/sizes + 4 (leaf K, internal K) + 4 flags
        off = 24
        off += 4 * 8  # base, free-space, EOF, driver-info addresses
        # root group symbol table entry
        self.root_header = r.u(off + 8, 8)
        self.r = r

    # ------------------------------------------------------------------ #
    def _messages(self, addr):
        r = self.r
        if self.buf[addr] != 1:
            raise NotImplementedError("object header version %d" % self.buf[addr])
        nmsg = r.u(addr + 2, 2)
        hsize = r.u(addr + 8, 4)
        blocks = [(addr + 16, hsize)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(msgs) < nmsg:
                mtype = r.u(p, 2)
                msize = r.u(p + 2, 2)
                data = p + 8
                if mtype == 0x10:  # continuation
                    blocks.append((r.u(data, 8), r.u(data + 8, 8)))
                msgs.append((mtype, data, msize))
                p = data + msize
        return msgs

    def _group_entries(self, header_addr):
        r = self.r
        btree = heap = None
        for mtype, data, _ in self._messages(header_addr):
            if mtype == 0x11:
                btree = r.u(data, 8)
                heap = r.u(data + 8, 8)
        if btree is None:
            raise H5Error("not a symbol-table group")
        if self.buf[heap:heap + 4] != b"HEAP":
            raise H5Error("bad local heap")
        heap_data = r.u(heap + 24, 8)
        out = {}
        self._walk_btree(btree, heap_data, out)
        return out

    def _walk_btree(self, addr, heap_data, out):
        r = self.r
        if self.buf[addr:addr + 4] != b"TREE":
            raise H5Error("bad B-tree node")
        level = self.buf[addr + 5]
        used = r.u(addr + 6, 2)
        p = addr + 8 + 16  # signature/type/level/used + siblings
        for i in range(used):
            p += 8  # key
            child = r.u(p, 8)
            p += 8
            if level > 0:
                self._walk_btree(child, heap_data, out)
            else:
                self._read_snod(child, heap_data, out)

    def _read_snod(self, addr, heap_data, out):
        r = self.r
        if self.buf[addr:addr + 4] != b"SNOD":
            raise H5Error("bad symbol node")
        nsym = r.u(addr + 6, 2)
        p = addr + 8
        for _ in range(nsym):
            name_off = r.u(p, 8)
            hdr = r.u(p + 8, 8)
            s = heap_data + name_off
            e = self.buf.index(b"\x00", s)
            out[self.buf[s:e].decode()] = hdr
            p += 40

    # ------------------------------------------------------------------ #
    def _datatype(self, data):
        r = self.r
        cls = self.buf[data] & 0x0F
        bits = r.u(data + 1, 3)
        size = r.u(data + 4, 4)
        if cls == 1:  # floating point
            if bits & 1:
                raise NotImplementedError("big-endian float")
            return np.dtype("<f%d" % size)
        if cls == 0:  # fixed point
            signed = bool(bits & 0x08)
            return np.dtype(("<i%d" if signed else "<u%d") % size)
        if cls == 3:  # fixed-length string
            return np.dtype("S%d" % size)
        if cls == 6:  # compound (version 1 members)
            nmemb = bits & 0xFFFF
            p = data + 8
            fields = []
            for _ in range(nmemb):
                e = self.buf.index(b"\x00", p)
                name = self.buf[p:e].decode()
                p += ((e - p) // 8 + 1) * 8
                moff = r.u(p, 4)
                p += 4 + 1 + 3 + 4 + 4 + 16  # offset, dim, rsvd, perm, rsvd, dims
                sub = self._datatype(p)
                p += 8 + self._props_len(p)
                fields.append((name, sub, moff))
            return np.dtype({"names": [f[0] for f in fields],
                             "formats": [f[1] for f in fields],
                             "offsets": [f[2] for f in fields],
                             "itemsize": size})
        raise NotImplementedError("datatype class %d" % cls)

    def _props_len(self, data):
        cls = self.buf[data] & 0x0F
        if cls == 1:
            return 12
        if cls == 0:
            return 4
        raise NotImplementedError("nested datatype class %d" % cls)

    def _dataspace(self, data):
        r = self.r
        ver = self.buf[data]
        ndim = self.buf[data + 1]
        if ver == 1:
            p = data + 8
        elif ver == 2:
            p = data + 4
        else:
            raise NotImplementedError("dataspace version %d" % ver)
        return tuple(r.u(p + 8 * i, 8) for i in range(ndim))

    def _layout(self, data):
        r = self.r
        ver = self.buf[data]
        if ver != 3:
            raise NotImplementedError("layout version %d" % ver)
        cls = self.buf[data + 1]
        if cls != 1:
            raise NotImplementedError("layout class %d" % cls)
        addr = r.u(data + 2, 8)
        size = r.u(data + 10, 8)
        return (None if addr == self.UNDEF else addr), size

    # ------------------------------------------------------------------ #
    def get(self, path):
        hdr = self.root_header
        for part in [p for p in path.split("/") if p]:
            entries = self._group_entries(hdr)
            if part not in entries:
                raise KeyError(path)
            hdr = entries[part]
        shape = dtype = None
        addr = size = None
        kind = "group"
        for mtype, data, _ in self._messages(hdr):
            if mtype == 0x01:
                shape = self._dataspace(data)
            elif mtype == 0x03:
                dtype = self._datatype(data)
            elif mtype == 0x08:
                addr, size = self._layout(data)
                kind = "dataset"
        if kind == "group":
            return sorted(self._group_entries(hdr).keys())
        return Dataset(self, path, shape, dtype, addr, size)

    def keys(self, path="/"):
        hdr = self.root_header
        for part in [p for p in path.split("/") if p]:
            hdr = self._group_entries(hdr)[part]
        return sorted(self._group_entries(hdr).keys())


def read(path, name):
    return File(path).get(name).read()


if __name__ == "__main__":
    import sys
    f = File(sys.argv[1])
    def walk(p):
        for k in f.keys(p):
            q = (p.rstrip("/") + "/" + k)
            obj = f.get(q)
            if isinstance(obj, Dataset):
                print(q, obj.shape, obj.dtype, obj.address, obj.size)
            else:
                walk(q)
    walk("/")
    _ = struct
