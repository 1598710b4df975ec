"""h5io reads the two run-file features that h5py writes for the reference
launcher (bfps/DNS.py:396-417, bfps/_base.py:158-174) and the library writer
does not use: chunked datasets (the extendable statistics series, created
with maxshape=None) and variable-length strings (h5py's str).

No file written by h5py/libhdf5 is available here (neither is installed and
the reference holds none with these features), so these readers are parity
unpinned: the fixture below is assembled byte by byte from the HDF5 file
format specification (data layout message v3 class 2, version-1 B-tree node
type 1 with chunk-size / filter-mask / offset keys, datatype class 9,
global heap collections), then read back through h5io.File."""
import struct

import numpy as np

from bfps_amd import h5io

UNDEF = 0xFFFFFFFFFFFFFFFF


def _space(shape, maxshape):
    """Dataspace message v1 with maximum dimensions."""
    rank = len(shape)
    return (struct.pack("<BBB5x", 1, rank, 1) + b"".join(struct.pack("<Q", s) for s in shape)
            + b"".join(struct.pack("<Q", m) for m in maxshape))


def _vlen_str_type():
    """Datatype class 9 (variable length), type string, ASCII, null padded;
    16-byte elements; base type one unsigned byte."""
    return bytes([0x19, 0x01, 0x00, 0x00]) + struct.pack("<I", 16) + h5io._encode_datatype(np.uint8)


class _Tail:
    """Appends raw structures at the end of a file."""

    def __init__(self, fh):
        fh.seek(0, 2)
        self.fh, self.pos = fh, (fh.tell() + 7) // 8 * 8

    def put(self, b):
        addr = self.pos
        self.fh.seek(addr)
        self.fh.write(b)
        self.pos = (addr + len(b) + 7) // 8 * 8
        return addr


def _entry_header_field(path, group, name):
    """File offset of the object-header address in the symbol-table entry
    of group/name."""
    p = h5io._Parser(path)
    try:
        g = p.object(p.root)
        for part in group.split("/"):
            g = g[part]
        heap = p.read(g.heap, 32)
        names = p.read(int.from_bytes(heap[24:32], "little"), int.from_bytes(heap[8:16], "little"))
        out = {}
        p._btree(g.btree, names, out)
        return out[name][1] - 16  # the entry's cached pair sits 24 bytes in, the header address 8 bytes in
    finally:
        p.f.close()


def _chunk_btree(tail, keys, children, level):
    """v1 B-tree node of type 1 (rank 3 data + the element "dimension")."""
    body = b"TREE" + bytes([1, level]) + struct.pack("<H", len(children)) + struct.pack("<QQ", UNDEF, UNDEF)
    for k, c in zip(keys, children):
        body += k + struct.pack("<Q", c)
    body += keys[len(children)]
    return tail.put(body)


def _key(nbytes, off):
    return struct.pack("<II", nbytes, 0) + b"".join(struct.pack("<Q", o) for o in tuple(off) + (0,))


def build_run_file(path):
    series = np.arange(5 * 3 * 4, dtype="<f8").reshape(5, 3, 4) * 0.5 - 3.0
    names = [b"NSVEparticles", b"", b"a longer variable-length string of 41 b."]
    with h5io.File(path, "w") as f:
        f["statistics/moments/velocity"] = np.zeros(1)   # replaced by a chunked dataset
        f["parameters/dns_type"] = "placeholder"         # replaced by a variable-length string
        f["parameters/names"] = np.zeros(1)              # replaced by an array of them
        f["parameters/niter_todo"] = np.int64(8)
    with open(path, "r+b") as fh:
        tail = _Tail(fh)
        # chunked [5][3][4] f64 in chunks of [2][3][4] (the last one partial),
        # maxshape (unlimited, 3, 4), under a two-level chunk B-tree
        cshape = (2, 3, 4)
        nbytes = int(np.prod(cshape)) * 8
        chunks = []
        for o in (0, 2, 4):
            c = np.zeros(cshape)
            part = series[o:o + 2]
            c[:len(part)] = part
            chunks.append(tail.put(c.tobytes()))
        leaf_a = _chunk_btree(tail, [_key(nbytes, (0, 0, 0)), _key(nbytes, (2, 0, 0)), _key(0, (4, 0, 0))],
                              chunks[:2], 0)
        leaf_b = _chunk_btree(tail, [_key(nbytes, (4, 0, 0)), _key(0, (6, 0, 0))], chunks[2:], 0)
        root = _chunk_btree(tail, [_key(nbytes, (0, 0, 0)), _key(nbytes, (4, 0, 0)), _key(0, (6, 0, 0))],
                            [leaf_a, leaf_b], 1)
        layout = bytes([3, 2, 4]) + struct.pack("<Q", root) + struct.pack("<IIII", *cshape, 8)
        hdr_series = tail.put(h5io._object_header([
            h5io._msg(0x01, _space(series.shape, (UNDEF, 3, 4))),
            h5io._msg(0x03, h5io._encode_datatype(np.float64), flags=1),
            h5io._msg(0x08, layout)]))
        # a global heap collection holding the strings (objects 1..3), then
        # its free space (object 0)
        body = b""
        for i, s in enumerate(names):
            body += struct.pack("<HH4xQ", i + 1, 1, len(s)) + s + b"\x00" * (-len(s) % 8)
        size = 4096
        free = size - 16 - len(body)
        coll = b"GCOL" + bytes([1, 0, 0, 0]) + struct.pack("<Q", size) + body + struct.pack("<HH4xQ", 0, 0, free)
        coll += b"\x00" * (size - len(coll))
        gcol = tail.put(coll)

        def vlen_dataset(elems, shape):
            raw = b"".join(struct.pack("<IQI", len(names[i]), gcol, i + 1) for i in elems)
            daddr = tail.put(raw)
            space = (struct.pack("<BBB5x", 1, 0, 0) if shape == () else _space(shape, shape))
            return tail.put(h5io._object_header([
                h5io._msg(0x01, space),
                h5io._msg(0x03, _vlen_str_type(), flags=1),
                h5io._msg(0x08, bytes([3, 1]) + struct.pack("<QQ", daddr, len(raw)), flags=1)]))

        hdr_name = vlen_dataset([0], ())
        hdr_names = vlen_dataset([0, 1, 2], (3,))
        fh.seek(40)
        fh.write(struct.pack("<Q", tail.pos))
    for group, name, hdr in (("statistics/moments", "velocity", hdr_series), ("parameters", "dns_type", hdr_name),
                             ("parameters", "names", hdr_names)):
        at = _entry_header_field(path, group, name)
        with open(path, "r+b") as fh:
            fh.seek(at)
            fh.write(struct.pack("<Q", hdr))
    return series, names


def test_chunked_and_vlen_reading(tmp_path):
    path = str(tmp_path / "run.h5")
    series, names = build_run_file(path)
    with h5io.File(path, "r") as f:
        got = f["statistics/moments/velocity"]
        assert got.shape == series.shape and np.array_equal(got, series)
        assert f["parameters/dns_type"] == names[0]
        assert list(f["parameters/names"]) == names
        assert int(f["parameters/niter_todo"]) == 8


def test_append_to_a_file_with_chunked_data(tmp_path):
    """The launcher appends to its run file: new datasets go to the end,
    the chunked and variable-length ones are kept as they are."""
    path = str(tmp_path / "run.h5")
    series, names = build_run_file(path)
    with h5io.File(path, "a") as f:
        f["statistics/moments/vorticity"] = np.ones((2, 3, 4))
        f["parameters/niter_todo"] = np.int64(16)
    with h5io.File(path, "r") as f:
        assert np.array_equal(f["statistics/moments/velocity"], series)
        assert np.array_equal(f["statistics/moments/vorticity"], np.ones((2, 3, 4)))
        assert f["parameters/dns_type"] == names[0]
        assert int(f["parameters/niter_todo"]) == 16


def test_filtered_chunks_are_refused(tmp_path):
    """Compressed (filtered) chunks are outside the subset: the reader says
    so instead of returning wrong data."""
    path = str(tmp_path / "run.h5")
    build_run_file(path)
    with open(path, "r+b") as fh:
        raw = bytearray(fh.read())
    # flip the filter mask of the first chunk key of the first leaf node
    i = raw.index(b"TREE\x01\x00")
    raw[i + 24 + 4] = 1
    with open(path, "wb") as fh:
        fh.write(raw)
    import pytest
    with h5io.File(path, "r") as f:
        with pytest.raises(h5io.H5FormatError):
            f["statistics/moments/velocity"]
