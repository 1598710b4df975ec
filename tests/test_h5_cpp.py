"""The C++ checkpoint-file layer (include/bfps_hip_h5.hpp, through
tests/cpp/h5_check) against the reference's own HDF5 checkpoint
(bfps/test/B32p1e4_checkpoint_0.h5, written by the HDF5 C library) and the
two Python readers (bfps_amd/h5io.py and the independent test reader
tests/golden/h5mini.py):

* every dataset of the reference file reads equal to the values h5mini
  extracted (B32p1e4.npz), in FP32 and FP64;
* datasets appended to a copy of the reference file (a new iteration in an
  existing group, new nested groups, more entries than one symbol-table
  node holds) read back equal through all three readers, the old datasets
  unchanged, and the file grows by the new bytes plus a few KB of group
  structure (nothing is rewritten);
* a dataset written again with the same type and shape is overwritten in
  place; another shape is refused;
* a fresh file reads back through the Python readers.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bfps_amd import h5io

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
REF = os.path.join(GOLD, "B32p1e4_checkpoint_0.h5")
TOOL_SRC = os.path.join(HERE, "cpp", "h5_check.cpp")
TOOL = os.path.join(HERE, "cpp", "h5_check")


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(TOOL) or os.path.getmtime(TOOL) < max(
            os.path.getmtime(TOOL_SRC), os.path.getmtime(os.path.join(ROOT, "include", "bfps_hip_h5.hpp"))):
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                               "-I" + os.path.join(ROOT, "include"), TOOL_SRC, "-o", TOOL])
    return TOOL


@pytest.fixture(scope="module")
def h5mini():
    import sys
    sys.path.insert(0, GOLD)
    import h5mini as m
    return m


def run(tool, *args):
    return subprocess.run([tool] + [str(a) for a in args], check=True, capture_output=True, text=True).stdout


def cread(tool, path, name, tmp_path):
    info = run(tool, "info", path, name).split()
    kind, shape = info[0], tuple(int(s) for s in info[2:])
    out = tmp_path / "out.raw"
    run(tool, "read", path, name, out)
    v = np.fromfile(out, dtype=np.float64)
    if kind == "c":
        return (v[0::2] + 1j * v[1::2]).reshape(shape)
    return v.reshape(shape)


def cwrite(tool, path, name, arr, tmp_path):
    arr = np.ascontiguousarray(arr)
    kind = {"f": "f", "i": "i", "c": "c"}[arr.dtype.kind]
    raw = tmp_path / "in.raw"
    arr.tofile(raw)
    run(tool, "write", path, name, kind, arr.dtype.itemsize, raw, *arr.shape)


def mini_read(m, path, name):
    a = m.read(path, name)
    if a.dtype.names == ("r", "i"):
        a = a["r"] + 1j * a["i"]
    return a


def test_reads_the_reference_file(tool, golden, tmp_path):
    assert run(tool, "keys", REF, "/").split() == ["tracers0", "vorticity"]
    assert run(tool, "keys", REF, "vorticity/complex").split() == ["0", "32", "64"]
    for it in (0, 32, 64):
        v = cread(tool, REF, "vorticity/complex/%d" % it, tmp_path)
        assert v.shape == (32, 32, 17, 3)
        assert np.array_equal(v, golden["vort_%d" % it].astype(np.complex128))
        assert np.array_equal(cread(tool, REF, "tracers0/state/%d" % it, tmp_path), golden["state_%d" % it])
        assert np.array_equal(cread(tool, REF, "tracers0/rhs/%d" % it, tmp_path), golden["rhs_%d" % it])


def test_append_in_place(tool, golden, h5mini, tmp_path):
    path = str(tmp_path / "ck.h5")
    shutil.copy(REF, path)
    size0 = os.path.getsize(path)
    rng = np.random.default_rng(3)
    w = (rng.standard_normal((32, 32, 17, 3)) + 1j * rng.standard_normal((32, 32, 17, 3))).astype(np.complex64)
    st = rng.standard_normal((10000, 3))
    cwrite(tool, path, "vorticity/complex/96", w, tmp_path)
    cwrite(tool, path, "tracers0/state/96", st, tmp_path)
    many = {"extra/deep/v%02d" % i: np.arange(i + 1, dtype=np.int64) for i in range(20)}
    for name, a in many.items():
        cwrite(tool, path, name, a, tmp_path)
    run(tool, "group", path, "empty/group")
    grown = os.path.getsize(path) - size0
    new_bytes = w.nbytes + st.nbytes + sum(a.nbytes for a in many.values())
    assert new_bytes <= grown < new_bytes + 64 * 1024
    # the three readers agree, old datasets unchanged
    f = h5io.File(path)
    for reader in (lambda n: cread(tool, path, n, tmp_path), lambda n: f[n], lambda n: mini_read(h5mini, path, n)):
        assert np.array_equal(reader("vorticity/complex/96"), w)
        assert np.array_equal(reader("tracers0/state/96"), st)
        assert np.array_equal(reader("vorticity/complex/32"), golden["vort_32"])
        assert np.array_equal(reader("tracers0/rhs/64"), golden["rhs_64"])
        for name, a in many.items():
            assert np.array_equal(reader(name), a)
    assert f.keys("vorticity/complex") == ["0", "32", "64", "96"]
    assert f.keys("empty/group") == []
    f.close()
    assert run(tool, "keys", path, "extra/deep").split() == sorted(n.split("/")[-1] for n in many)


def test_overwrite_in_place_and_refusals(tool, tmp_path):
    path = str(tmp_path / "ow.h5")
    run(tool, "fresh", path)
    a = np.arange(12, dtype=np.float64).reshape(3, 4)
    cwrite(tool, path, "g/a", a, tmp_path)
    size = os.path.getsize(path)
    cwrite(tool, path, "g/a", -a, tmp_path)
    assert os.path.getsize(path) == size
    assert np.array_equal(cread(tool, path, "g/a", tmp_path), -a)
    with pytest.raises(subprocess.CalledProcessError):
        cwrite(tool, path, "g/a", np.zeros(5), tmp_path)
    with pytest.raises(subprocess.CalledProcessError):
        cwrite(tool, path, "g", np.zeros(5), tmp_path)
    with pytest.raises(subprocess.CalledProcessError):
        run(tool, "read", path, "g/missing", tmp_path / "x")


def test_fresh_file_reads_in_python(tool, h5mini, tmp_path):
    path = str(tmp_path / "fresh.h5")
    run(tool, "fresh", path)
    v = (np.arange(6) + 1j * np.arange(6)[::-1]).astype(np.complex128).reshape(2, 3)
    s = np.linspace(0, 1, 7, dtype=np.float32)
    cwrite(tool, path, "vorticity/complex/0", v, tmp_path)
    cwrite(tool, path, "x", s, tmp_path)
    f = h5io.File(path)
    assert np.array_equal(f["vorticity/complex/0"], v) and f["x"].dtype == np.float32
    assert np.array_equal(f["x"], s)
    f.close()
    assert np.array_equal(mini_read(h5mini, path, "vorticity/complex/0"), v)
