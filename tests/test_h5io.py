"""The checkpoint-file layer (bfps_amd/h5io.py) against the reference's own
HDF5 checkpoint (bfps/test/B32p1e4_checkpoint_0.h5, written by the HDF5 C
library and h5py) and against the independent test reader
tests/golden/h5mini.py:

* every dataset of the reference file reads equal to the values h5mini
  extracted (B32p1e4.npz);
* files written here read back equal through both readers, for every
  supported type, nested and empty groups, and groups large enough for a
  multi-level B-tree;
* the object headers written for a complex field and a tracer array carry
  the same messages, flags and type encodings as the reference file's;
* appending to the reference file keeps its datasets and adds the new one.
"""
import os
import shutil

import numpy as np
import pytest

from bfps_amd import h5io

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = os.path.join(GOLD, "B32p1e4_checkpoint_0.h5")


@pytest.fixture(scope="module")
def h5mini():
    import sys
    sys.path.insert(0, GOLD)
    import h5mini as m
    return m


def mini_read(m, path, name):
    a = m.read(path, name)
    if a.dtype.names == ("r", "i"):
        a = a["r"] + 1j * a["i"]
    return a


def test_read_reference_checkpoint(golden):
    with h5io.File(REF) as f:
        assert f.keys() == ["tracers0", "vorticity"]
        assert f.keys("vorticity/complex") == ["0", "32", "64"]
        assert f.keys("tracers0") == ["rhs", "state"]
        for it in (0, 32, 64):
            v = f["vorticity/complex/%d" % it]
            assert v.dtype == np.complex64 and v.shape == (32, 32, 17, 3)
            assert np.array_equal(v, golden["vort_%d" % it])
            assert np.array_equal(f["tracers0/state/%d" % it], golden["state_%d" % it])
            assert np.array_equal(f["tracers0/rhs/%d" % it], golden["rhs_%d" % it])
        assert "vorticity/complex/32" in f and "vorticity/complex/33" not in f


def test_roundtrip_both_readers(tmp_path, h5mini):
    rng = np.random.default_rng(3)
    arrays = {
        "vorticity/complex/0": (rng.standard_normal((8, 8, 5, 3)) + 1j * rng.standard_normal((8, 8, 5, 3))).astype(np.complex64),
        "vorticity/complex/8": rng.standard_normal((8, 8, 5, 3)) + 1j * rng.standard_normal((8, 8, 5, 3)),
        "tracers0/state/0": rng.random((100, 3)),
        "tracers0/rhs/0": rng.random((4, 100, 3)),
        "a/b/c/f32": rng.standard_normal(17).astype(np.float32),
        "a/i64": np.arange(-5, 7, dtype=np.int64),
        "scalar": np.float64(2.5),
    }
    p = str(tmp_path / "t.h5")
    with h5io.File(p, "w") as f:
        for k, v in arrays.items():
            f[k] = v
        f.create_group("empty/inner")
    with h5io.File(p) as f:
        for k, v in arrays.items():
            got = f[k]
            assert got.dtype == np.asarray(v).dtype and np.array_equal(got, v), k
        assert f.keys("empty/inner") == []
    for k, v in arrays.items():
        if np.ndim(v):  # the test reader has no scalar dataspaces
            assert np.array_equal(mini_read(h5mini, p, k), v), k


def test_headers_match_reference(tmp_path, h5mini):
    p = str(tmp_path / "t.h5")
    with h5io.File(p, "w") as f:
        f["vorticity/complex/0"] = np.zeros((32, 32, 17, 3), np.complex64)
        f["tracers0/state/0"] = np.zeros((10000, 3))
    mine, ref = h5mini.File(p), h5mini.File(REF)

    def msgs(F, path):
        addr = F.root_header
        for q in path.split("/"):
            addr = F._group_entries(addr)[q]
        return [(t, F.buf[d - 4], bytes(F.buf[d:d + s])) for t, d, s in F._messages(addr) if t != 0]

    for rp, mp in (("vorticity/complex/32", "vorticity/complex/0"), ("tracers0/state/32", "tracers0/state/0")):
        a, b = msgs(ref, rp), msgs(mine, mp)
        assert [(t, fl) for t, fl, _ in a] == [(t, fl) for t, fl, _ in b]
        for (t, _, da), (_, _, db) in zip(a, b):
            if t in (0x01, 0x03, 0x05):  # dataspace, datatype, fill value: same bytes
                assert da == db, hex(t)
            else:  # layout (own address), modification time: same version and class
                assert da[:2] == db[:2], hex(t)


def test_large_group_multilevel_btree(tmp_path, h5mini):
    p = str(tmp_path / "t.h5")
    n = 600  # > 2 LEAF_K x 2 INTERNAL_K entries: two B-tree levels
    with h5io.File(p, "w") as f:
        for i in range(n):
            f["g/%d" % i] = np.array([i, -i], dtype=np.int64)
    with h5io.File(p) as f:
        names = f.keys("g")
        assert sorted(names) == sorted(str(i) for i in range(n))
        for i in (0, 7, 8, 255, 256, 599):
            assert list(f["g/%d" % i]) == [i, -i]
    assert sorted(h5mini.File(p).keys("/g")) == sorted(str(i) for i in range(n))
    assert list(h5mini.read(p, "g/513")) == [513, -513]


def test_append_to_reference_file(tmp_path, golden, h5mini):
    p = str(tmp_path / "B32p1e4_checkpoint_0.h5")
    shutil.copy(REF, p)
    extra = np.ones((32, 32, 17, 3), np.complex64) * (1 - 2j)
    with h5io.File(p, "a") as f:
        f["vorticity/complex/96"] = extra
    with h5io.File(p) as f:
        assert f.keys("vorticity/complex") == ["0", "32", "64", "96"]
        assert np.array_equal(f["vorticity/complex/96"], extra)
        for it in (0, 32, 64):
            assert np.array_equal(f["vorticity/complex/%d" % it], golden["vort_%d" % it])
            assert np.array_equal(f["tracers0/rhs/%d" % it], golden["rhs_%d" % it])
    assert np.array_equal(mini_read(h5mini, p, "vorticity/complex/96"), extra)
    assert np.array_equal(mini_read(h5mini, p, "tracers0/state/64"), golden["state_64"])


def test_errors(tmp_path):
    p = str(tmp_path / "x.h5")
    with open(p, "wb") as fh:
        fh.write(b"not hdf5" * 20)
    with pytest.raises(h5io.H5FormatError):
        h5io.File(p)
    with h5io.File(REF) as f:
        with pytest.raises(KeyError):
            f["vorticity/complex/5"]
        with pytest.raises(KeyError):
            f["vorticity/complex"]
        with pytest.raises(IOError):
            f["new"] = np.zeros(3)
    with pytest.raises(h5io.H5FormatError):
        with h5io.File(str(tmp_path / "y.h5"), "w") as f:
            f["s"] = np.array(["text"])


def test_append_costs_only_the_new_bytes(tmp_path, h5mini):
    """Mode "a" on an existing file appends: the file grows by the new data
    plus a few KB of group structure, a dataset of the same type and shape
    is overwritten in place (no growth), and the reference file's datasets,
    the new ones and the replaced one read equal through both readers."""
    path = str(tmp_path / "ck.h5")
    shutil.copy(REF, path)
    size0 = os.path.getsize(path)
    w = (np.arange(32 * 32 * 17 * 3) * (1 + 0.5j)).astype(np.complex64).reshape(32, 32, 17, 3)
    st = np.linspace(0, 1, 30000).reshape(10000, 3)
    with h5io.File(path, "a") as f:
        f["vorticity/complex/96"] = w
        f["tracers0/state/96"] = st
        f["new/deep/x"] = np.arange(5)
    grown = os.path.getsize(path) - size0
    assert w.nbytes + st.nbytes <= grown < w.nbytes + st.nbytes + 64 * 1024
    size1 = os.path.getsize(path)
    with h5io.File(path, "a") as f:
        f["tracers0/state/96"] = -st  # same type and shape: in place
    assert os.path.getsize(path) == size1
    with h5io.File(path, "a") as f:
        f["new/deep/x"] = np.arange(7)  # another shape: a new dataset
    g = np.load(os.path.join(GOLD, "B32p1e4.npz"))
    with h5io.File(path) as f:
        for read in (lambda n: f[n], lambda n: mini_read(h5mini, path, n)):
            assert np.array_equal(read("vorticity/complex/96"), w)
            assert np.array_equal(read("tracers0/state/96"), -st)
            assert np.array_equal(read("new/deep/x"), np.arange(7))
            assert np.array_equal(read("vorticity/complex/64"), g["vort_64"])
            assert np.array_equal(read("tracers0/rhs/32"), g["rhs_32"])
        assert f.keys("vorticity/complex") == ["0", "32", "64", "96"]
