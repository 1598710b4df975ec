"""GPU test of the allocation-free cospectrum entry point
(bfpsh_field_cospectrum_noalloc, no counterpart in the reference): through
the C ABI it must give the shell sums of bfpsh_field_cospectrum (the same
kernel; the FP64 atomics make the summation order, and so the last bits,
differ: relative 1e-13 of the largest sum) on two calls in a row, the
second on a smaller grid (fewer shells, so the fixed scratch must be
cleared of the first call's sums), and reject a missing field as the
allocating entry point does."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    import bfps_amd
    bfps_amd.capi.load()
    return bfps_amd


def spectral_field(n, seed):
    rng = np.random.default_rng(seed)
    shp = (n, n, n // 2 + 1, 3)
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    K = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kx[None, None, :] ** 2)
    amp = (1.0 / (1.0 + K)) * (K > 0)
    return (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * amp[..., None]


def both(B, lib, fs, fa, fb):
    nsh = fs.kk.nshells
    ref = np.zeros((nsh, 3, 3))
    got = np.full((nsh, 3, 3), np.nan)
    B.capi.check(fs.handle, lib.bfpsh_field_cospectrum(fs.handle, fa.id, fb.id, ref.ctypes.data_as(B.capi._DP)),
                 "field_cospectrum")
    B.capi.check(fs.handle, lib.bfpsh_field_cospectrum_noalloc(fs.handle, fa.id, fb.id,
                                                                got.ctypes.data_as(B.capi._DP)),
                 "field_cospectrum_noalloc")
    return got, ref


@pytest.mark.parametrize("precision", ["double", "single"])
def test_noalloc_cospectrum_matches_allocating(B, precision):
    lib = B.capi.load()
    results = []
    for n, seed in ((64, 11), (32, 12)):  # the second call has fewer shells
        say("case: cospectrum_noalloc %s n=%d" % (precision, n))
        fs = B.vorticity_equation("na%d" % n, n, precision=precision)
        fs.io_checkpoint(read=True, data=spectral_field(n, seed).astype(fs.cdtype))
        fu, fw = B.field(fs, 3), B.field(fs, 3)
        fu.copy_from("velocity")
        fw.copy_from("vorticity")
        for fa, fb in ((fu, fu), (fu, fw), (fw, fw)):
            got, ref = both(B, lib, fs, fa, fb)
            assert np.all(np.isfinite(got))
            err = np.abs(got - ref).max() / np.abs(ref).max()
            say("  n=%d shells=%d rel diff %.3e" % (n, fs.kk.nshells, err))
            assert err <= 1e-13
        results.append(fs.kk.nshells)
        # a bad field id is rejected, as by the allocating entry point
        out = np.zeros((fs.kk.nshells, 3, 3))
        assert lib.bfpsh_field_cospectrum_noalloc(fs.handle, fu.id, 9999, out.ctypes.data_as(B.capi._DP)) != 0
        assert lib.bfpsh_field_cospectrum(fs.handle, fu.id, 9999, out.ctypes.data_as(B.capi._DP)) != 0
        fu.close()
        fw.close()
        fs.close()
    assert results[1] < results[0]
