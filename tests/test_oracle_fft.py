"""Pin the oracle's 3D real FFT (own Stockham code) against numpy.fft (pocketfft).

Layouts follow FFTW-MPI TRANSPOSED_OUT/IN as used by the reference
(bfps/cpp/field.cpp:72-118): real [z][y][x][c] <-> complex [ky][kz][kx][c],
unnormalised in both directions.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("shape", [(8, 8, 8, 3), (6, 6, 6, 1), (12, 8, 10, 2), (32, 32, 32, 3),
                                   (48, 48, 48, 1), (4, 4, 4, 9), (20, 20, 20, 3)])
def test_r2c_c2r_vs_numpy(oracle, shape):
    nx, ny, nz, nc = shape
    rng = np.random.default_rng(1)
    r = rng.standard_normal((nz, ny, nx, nc))
    c = oracle.fft3d_r2c(r)
    ref = np.fft.rfftn(r, axes=(0, 1, 2)).transpose(1, 0, 2, 3)
    assert np.abs(c - ref).max() <= 1e-14 * np.abs(ref).max()
    back = oracle.fft3d_c2r(c, nx)
    assert np.abs(back / (nx * ny * nz) - r).max() < 1e-13
    c32 = oracle.fft3d_r2c(r, fp32=True)
    assert np.abs(c32 - ref).max() <= 1e-6 * np.abs(ref).max()


def test_c2r_ignores_imag_of_dc_and_nyquist(oracle):
    """c2r semantics: imaginary parts of kx=0 / kx=N/2 after the (ky,kz)
    transforms are dropped (Hermitian projection), as FFTW's hc2r does."""
    n = 8
    rng = np.random.default_rng(2)
    c = rng.standard_normal((n, n, n // 2 + 1, 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1, 1))
    out = oracle.fft3d_c2r(c, n)
    # reference: numpy irfftn applies the same convention along the last axis
    ref = np.fft.irfftn(c.transpose(1, 0, 2, 3), s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert np.abs(out - ref).max() < 1e-12


def test_brute_force_dft_tiny(oracle):
    n = 4
    rng = np.random.default_rng(3)
    r = rng.standard_normal((n, n, n, 1))
    c = oracle.fft3d_r2c(r)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for ky in range(n):
        for kz in range(n):
            for kx in range(n // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * x + ky * y + kz * z) / n)
                assert abs((r[..., 0] * ph).sum() - c[ky, kz, kx, 0]) < 1e-12


def test_symmetrize_pins(oracle):
    """field::symmetrize (bfps/cpp/field.cpp:900-978): identity on the r2c of a
    real field; on arbitrary data it makes the kx = 0 plane Hermitian for
    every mode it writes and touches nothing with kx > 0."""
    rng = np.random.default_rng(4)
    for (nx, ny, nz, nc) in [(8, 8, 8, 3), (12, 10, 6, 1), (16, 16, 16, 9)]:
        r = rng.standard_normal((nz, ny, nx, nc))
        c = np.fft.rfftn(r, axes=(0, 1, 2)).transpose(1, 0, 2, 3)
        s = oracle.symmetrize(c, nx)
        assert np.abs(s - c).max() <= 1e-13 * np.abs(c).max()
        a = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
        s = oracle.symmetrize(a, nx)
        assert np.array_equal(s[:, :, 1:], a[:, :, 1:])
        assert s[0, 0, 0].imag.max() == 0 and np.array_equal(s[0, 0, 0].real, a[0, 0, 0].real)
        for ky in range(ny):
            for kz in range(nz):
                my, mz = (-ky) % ny, (-kz) % nz
                if ky > ny // 2 or (ky == 0 and kz > nz // 2):
                    # written: conjugate of the mirror, which is never written
                    assert np.array_equal(s[ky, kz, 0], np.conj(a[my, mz, 0]))
                    assert np.array_equal(s[my, mz, 0], a[my, mz, 0])
                elif not (ky == 0 and kz == 0):
                    assert np.array_equal(s[ky, kz, 0], a[ky, kz, 0])
