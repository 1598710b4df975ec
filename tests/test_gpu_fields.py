"""GPU parity of the standalone field transforms (field::dft / ift /
symmetrize, bfps/cpp/field.cpp:144-162, 900-978) against the CPU oracle.

The device field does not hold the kx = N/2 plane (it lies outside every
dealiasing sphere of the solver), so the oracle's transforms are compared
with that plane zeroed on the spectral side.

Tolerances: FP64 relative L2 <= 1e-13 (rounding of an N log N transform);
FP32 relative L2 <= 1e-6 against the FP64 oracle and within 3x the FP32
oracle's own error.
"""
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


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def drop_nyquist(c):
    c = c.copy()
    c[:, :, -1] = 0
    return c


CASES = [(32, 1, "double"), (32, 3, "single"), (48, 9, "double"), (48, 3, "single"),
         (64, 3, "double"), (64, 9, "single"), (128, 1, "single"), (128, 3, "double")]


@pytest.mark.parametrize("n,nc,precision", CASES)
def test_dft_vs_oracle(B, oracle, n, nc, precision):
    say("case: field dft n=%d nc=%d %s" % (n, nc, precision))
    rng = np.random.default_rng(n + nc)
    r = rng.standard_normal((n, n, n, nc))
    fs = B.vorticity_equation("t", n, precision=precision)
    f = B.field(fs, nc)
    f.put(r, real_space=True)
    f.dft()
    got = f.get()
    ref = drop_nyquist(oracle.fft3d_r2c(r))
    e = rel(got, ref)
    say("  rel", e)
    if precision == "double":
        assert e < 1e-13
    else:
        e32 = rel(drop_nyquist(oracle.fft3d_r2c(r.astype(np.float32), fp32=True)), ref)
        say("  oracle fp32 rel", e32)
        assert e < 1e-6 and e < 3 * e32 + 1e-8


@pytest.mark.parametrize("n,nc,precision", CASES)
def test_ift_vs_oracle(B, oracle, n, nc, precision):
    """Arbitrary complex input (not Hermitian on kx = 0): c2r keeps the real
    part of each kx = 0 coefficient after the y and z transforms, as the
    oracle (and FFTW's c2r) does."""
    say("case: field ift n=%d nc=%d %s" % (n, nc, precision))
    rng = np.random.default_rng(100 + n + nc)
    shp = (n, n, n // 2 + 1, nc)
    c = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    fs = B.vorticity_equation("t", n, precision=precision)
    f = B.field(fs, nc)
    f.put(c, real_space=False)
    f.ift()
    assert f.real_space_representation
    got = f.get()
    ref = oracle.fft3d_c2r(drop_nyquist(c), n)
    e = rel(got, ref)
    say("  rel", e)
    assert e < (1e-13 if precision == "double" else 1e-6)


@pytest.mark.parametrize("n,precision", [(256, "single"), (128, "double")])
def test_round_trip_large(B, n, precision):
    """ift(dft(r)) = N^3 r for r band-limited below the x-Nyquist mode;
    checked on 3 components at full size without a CPU transform."""
    say("case: field round trip n=%d %s" % (n, precision))
    rng = np.random.default_rng(7)
    r = rng.standard_normal((n, n, n, 3)).astype(np.float32 if precision == "single" else np.float64)
    fs = B.vorticity_equation("t", n, precision=precision)
    f = B.field(fs, 3)
    # band-limit on the device: dft, drop the kx = N/2 plane (not stored), ift
    f.put(r, real_space=True)
    f.dft()
    f.ift()
    r1 = f.get() / n ** 3
    f.put(r1, real_space=True)
    f.dft()
    f.ift()
    r2 = f.get() / n ** 3
    e = rel(r2, r1)
    say("  rel", e)
    assert e < (1e-14 if precision == "double" else 1e-6)
    # removing the x-Nyquist mode changes r by its (small) share of the energy
    assert 0.0 < rel(r1, r) < 0.2


@pytest.mark.parametrize("n,nc,precision", [(32, 3, "double"), (48, 9, "single"), (64, 1, "double")])
def test_symmetrize_vs_oracle(B, oracle, n, nc, precision):
    say("case: field symmetrize n=%d nc=%d %s" % (n, nc, precision))
    rng = np.random.default_rng(11)
    shp = (n, n, n // 2 + 1, nc)
    c = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    if precision == "single":
        c = c.astype(np.complex64)
    fs = B.vorticity_equation("t", n, precision=precision)
    f = B.field(fs, nc)
    f.put(c, real_space=False)
    f.symmetrize()
    got = f.get()
    ref = drop_nyquist(oracle.symmetrize(c, n))
    assert np.array_equal(got.astype(np.complex128), ref)


def test_copy_from_and_cospectrum(B, oracle):
    n = 32
    say("case: field copy_from / cospectrum")
    rng = np.random.default_rng(3)
    shp = (n, n, n // 2 + 1, 3)
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    K = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kx[None, None, :] ** 2)
    w = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * (np.exp(-(K / 3) ** 2) * (K > 0))[..., None]
    fs = B.vorticity_equation("t", n, precision="double")
    fs.io_checkpoint(read=True, data=w)
    fu = B.field(fs, 3)
    fw = B.field(fs, "THREE")
    fu.copy_from("velocity")
    fw.copy_from("vorticity")
    assert np.array_equal(fw.get(), fs.cvorticity)
    assert np.array_equal(fu.get(), fs.compute_velocity())
    for a, b, fa, fb in [("velocity", "velocity", fu, fu), ("velocity", "vorticity", fu, fw)]:
        # shell sums use atomics: equal up to summation order
        g, r = fa.cospectrum(fb), fs.cospectrum(a, b)
        assert np.abs(g - r).max() <= 1e-13 * np.abs(r).max()
    o = oracle.Solver(n, precision="double")
    o.set_cfield("vorticity", w)
    o.io_projection()
    o.compute_velocity()
    r = o.cospectrum("velocity", "vorticity")
    assert np.abs(fu.cospectrum(fw) - r).max() <= 1e-12 * np.abs(r).max()
    # real-space velocity through the generic transform equals the solver's
    fu.ift()
    fs.velocity_rspace()
    assert rel(fu.get(), fs.get_rfield(B.capi.RVELOCITY)) < 1e-14


def test_field_errors(B):
    fs = B.vorticity_equation("t", 16, precision="double")
    with pytest.raises(B.capi.BfpshError):
        B.field(fs, 2)
    f = B.field(fs, 1)
    with pytest.raises(B.capi.BfpshError):
        f.copy_from("velocity")
    i = f.id
    f.close()
    lib = B.capi.load()
    assert lib.bfpsh_field_ift(fs.handle, i) != 0
    assert lib.bfpsh_field_destroy(fs.handle, i) != 0
    assert lib.bfpsh_field_dft(fs.handle, 12345) != 0


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sample_standalone_fields(B, precision):
    """sample_compute_field on standalone fields equals the tracer sampling
    of the solver's own real-space velocity; a one-component float field is
    left unsampled as in the reference."""
    n, np_ = 32, 3000
    say("case: sample standalone fields %s" % precision)
    from bfps_amd import ic
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=4, k0=3.0))
    ps = B.particles_system(fs, np_, 3, 1, 4)
    ps.put(ic.uniform_positions(np_, seed=5))
    fs.velocity_rspace()
    ref = ps.sample("velocity")
    fu = B.field(fs, 3)
    fu.copy_from("velocity")
    fu.ift()
    got = ps.sample_compute_field(fu)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max() + (1e-6 if precision == "single" else 0)
    with pytest.raises(B.capi.BfpshError):
        fu.dft()
        ps.sample_compute_field(fu)
    f1 = B.field(fs, 1)
    f1.put(fs.get_rfield(B.capi.RVELOCITY)[..., :1], real_space=True)
    out = np.full((np_, 1), 7.0)
    ps.sample_compute_field(f1, out)
    if precision == "single":
        assert np.all(out == 7.0)
    else:
        assert np.abs(out[:, 0] - ref[:, 0]).max() <= 1e-12 * np.abs(ref).max()


# ------------------------------------------------------------ pass shapes (autotuner)
def _set_all(fs, idx):
    for k in ("z_inv", "y_inv", "z_fwd", "y_fwd"):
        fs.set_pass_shape(k, idx)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_every_pass_shape_vs_oracle_256(B, oracle, precision):
    """Every compiled tile shape of the strided passes gives the transforms
    of the oracle at 256^3 (one component): dft and ift within the
    tolerances of the default shape's tests."""
    n = 256
    rng = np.random.default_rng(256)
    r = rng.standard_normal((n, n, n, 1))
    ref = drop_nyquist(oracle.fft3d_r2c(r))
    back = oracle.fft3d_c2r(ref, n)
    fs = B.vorticity_equation("t", n, precision=precision, plan_rigor="FFTW_ESTIMATE")
    nsh = fs.pass_shapes()["y_inv"][3]
    say("case: %d pass shapes at 256^3 %s" % (nsh, precision))
    assert nsh > 1
    f = B.field(fs, 1)
    tol = 1e-13 if precision == "double" else 1e-6
    for idx in range(nsh):
        _set_all(fs, idx)
        e, b, _, _, pp = fs.pass_shapes()["y_inv"]
        f.put(r, real_space=True)
        f.dft()
        c = f.get()
        e_f = rel(c, ref)
        f.ift()
        e_b = rel(f.get(), back)
        say("  shape %d (E=%d, B=%d, persistent %d): dft rel %.3e, ift(dft) rel %.3e" % (idx, e, b, pp, e_f, e_b))
        assert e_f < tol and e_b < tol
    f.close()
    fs.close()


@pytest.mark.parametrize("n,precision", [(512, "single"), (512, "double"), (1024, "single")])
def test_every_pass_shape_matches_default(B, n, precision):
    """At 512^3 and 1024^3 every shape's dft and ift equal the built-in
    shape's (validated against the oracle by the step and transform tests)
    within rounding: FP64 1e-13, FP32 1e-6 relative L2."""
    rng = np.random.default_rng(n)
    dt = np.float32 if precision == "single" else np.float64
    r = rng.standard_normal((n, n, n, 1), dtype=dt)
    fs = B.vorticity_equation("t", n, precision=precision, plan_rigor="FFTW_ESTIMATE")
    nsh = fs.pass_shapes()["y_inv"][3]
    say("case: %d pass shapes at %d^3 %s" % (nsh, n, precision))
    f = B.field(fs, 1)
    tol = 1e-13 if precision == "double" else 1e-6
    c0 = r0 = None
    for idx in range(nsh):
        _set_all(fs, idx)
        e, b, _, _, pp = fs.pass_shapes()["y_inv"]
        f.put(r, real_space=True)
        f.dft()
        c = f.get()
        f.ift()
        rr = f.get()
        if idx == 0:
            c0, r0 = c, rr
            say("  shape 0 (E=%d, B=%d, persistent %d): reference" % (e, b, pp))
            continue
        e_f, e_b = rel(c, c0), rel(rr, r0)
        say("  shape %d (E=%d, B=%d, persistent %d): dft rel %.3e, ift rel %.3e" % (idx, e, b, pp, e_f, e_b))
        assert e_f < tol and e_b < tol
        del c, rr
    f.close()
    fs.close()


def test_plan_rigor_and_wisdom(B, tmp_path):
    """MEASURE picks a compiled shape per pass kind and records it in the
    wisdom; a new context of the same size with ESTIMATE then takes the
    wisdom's shapes; the exported wisdom file reads back."""
    B.wisdom_forget()
    fs = B.vorticity_equation("t", 256, precision="single", plan_rigor="FFTW_MEASURE")
    got = fs.pass_shapes()
    say("case: plan rigor at 256^3: %s" % got)
    for k, (e, b, idx, nsh, pp) in got.items():
        assert 0 <= idx < nsh and e > 0 and b > 0 and pp in (0, 1)
    path = str(tmp_path / "w.txt")
    B.wisdom_export(path)
    lines = [l for l in open(path) if not l.startswith("#")]
    assert len(lines) == 4 and all("/f32/256/" in l for l in lines)
    B.wisdom_forget()
    fs2 = B.vorticity_equation("t2", 256, precision="single", plan_rigor="FFTW_ESTIMATE")
    assert all(v[2] == 0 for v in fs2.pass_shapes().values())  # no wisdom: built-in shapes
    fs2.close()
    B.wisdom_import(path)
    fs3 = B.vorticity_equation("t3", 256, precision="single", plan_rigor="FFTW_ESTIMATE")
    assert fs3.pass_shapes() == got
    # the step runs with the tuned shapes (the step tests cover its values)
    fs3.io_checkpoint(read=True, data=np.zeros((256, 256, 129, 3), dtype=np.complex64))
    fs3.step(1e-3)
    assert np.abs(fs3.cvorticity).max() == 0.0
    fs3.close()
    fs.close()
    with pytest.raises(B.capi.BfpshError):
        fs4 = B.vorticity_equation("t4", 32, precision="single", plan_rigor="FFTW_WHATEVER")
        fs4.close()
    B.wisdom_forget()
