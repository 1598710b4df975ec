"""Pin the tracer interpolation weights.

Three independent constructions must agree:
  * the reference's own generated sources bfps/cpp/spline_n{1..10}.cpp and
    bfps/cpp/Lagrange_polys.cpp, compiled into oracle/_ref by
    oracle/build_ref.sh (skipped when that library is absent);
  * the oracle's numeric construction (Lagrange basis + two-point Hermite solve);
  * the committed table bfps_amd/csrc/interp_tables.h (exact rationals from
    bfps_amd/csrc/gen_interp_tables.py).
Plus mathematical invariants: partition of unity, reproduction of
polynomials up to the kernel's order, derivative consistency.
Kernel selection follows bfps/cpp/particles/particles_generic_interp.hpp:10-313
(m = 0 -> beta_Lagrange_n{n}, m >= 1 -> beta_n{n}_m{m}).
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libref_interp.so")
KERNELS = [(n, m) for n in range(1, 11) for m in range(3)]
XS = np.linspace(0, 1, 97)[:-1]


def header_table():
    src = open(os.path.join(ROOT, "bfps_amd", "csrc", "interp_tables.h")).read()
    off = [int(v) for v in re.search(r"offset\[30\] = \{\s*([^}]*)\}", src).group(1).split(",")]
    deg = [int(v) for v in re.search(r"degree\[30\] = \{\s*([^}]*)\}", src).group(1).split(",")]
    body = re.search(r"coef\[\d+\] = \{([^}]*)\}", src).group(1)
    coef = np.array([float(v) for v in body.replace("\n", "").split(",") if v.strip()])
    return off, deg, coef


def ev(tab, d, x):
    return np.array([np.polyval(tab[d, j, ::-1], x) for j in range(tab.shape[1])])


@pytest.fixture(scope="module")
def htab():
    return header_table()


@pytest.mark.parametrize("n,m", KERNELS)
def test_header_matches_oracle_construction(oracle, htab, n, m):
    off, deg, coef = htab
    k = (n - 1) * 3 + m
    W = 2 * n + 2
    tab = coef[off[k]:off[k] + 3 * W * (deg[k] + 1)].reshape(3, W, deg[k] + 1)
    otab = oracle.interp_table(n, m)
    assert otab.shape == tab.shape
    for d in range(3):
        for x in XS[::7]:
            a, b = ev(tab, d, x), ev(otab, d, x)
            assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(a).max())


@pytest.mark.parametrize("n,m", KERNELS)
def test_invariants(htab, n, m):
    off, deg, coef = htab
    k = (n - 1) * 3 + m
    W = 2 * n + 2
    tab = coef[off[k]:off[k] + 3 * W * (deg[k] + 1)].reshape(3, W, deg[k] + 1)
    nodes = np.arange(-n, n + 2)
    order = 2 * n + 1 if m == 0 else min(2 * n, 2 * m + 1)
    for x in XS:
        w = ev(tab, 0, x)
        assert abs(w.sum() - 1) < 1e-12
        assert abs(ev(tab, 1, x).sum()) < 1e-10
        # reproduces x itself (and higher powers up to the kernel's exactness)
        assert abs((w * nodes).sum() - x) < 1e-11
        if m == 0:
            for p in range(2, order + 1):
                assert abs((w * nodes.astype(float) ** p).sum() - x ** p) < 1e-7 * (n + 1) ** p


@pytest.mark.skipif(not os.path.exists(REFLIB), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("n,m", KERNELS)
def test_header_matches_reference_sources(htab, n, m):
    lib = ctypes.CDLL(REFLIB)
    fname = ("beta_Lagrange_n%d" % n) if m == 0 else ("beta_n%d_m%d" % (n, m))
    f = getattr(lib, "_Z%d%sidPd" % (len(fname), fname))
    f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    off, deg, coef = htab
    k = (n - 1) * 3 + m
    W = 2 * n + 2
    tab = coef[off[k]:off[k] + 3 * W * (deg[k] + 1)].reshape(3, W, deg[k] + 1)
    for d in range(3):
        for x in XS:
            out = (ctypes.c_double * W)()
            f(d, float(x), out)
            ref = np.array(out[:])
            assert np.abs(ref - ev(tab, d, x)).max() <= 1e-13 * max(1.0, np.abs(ref).max())
