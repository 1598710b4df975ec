"""GPU parity of the statistics either side of the time step against the CPU
oracle: compute_pressure and compute_Lagrangian_acceleration
(bfps/cpp/vorticity_equation.cpp:441-578), field::compute_rspace_stats
(bfps/cpp/field.cpp:661-850) and field::compute_stats (field.cpp:980-1040).

Tolerances: pressure and acceleration FP64 relative L2 <= 1e-12, FP32 <=
2e-5 against the FP64 oracle (both then agree with the ABC closed forms,
p = -|u|^2/2 + const and a = grad |u|^2/2 - nu u, to the same bounds).
Statistics of identical real data: histograms bit-exact, moments relative
1e-12 (only the order of the sums differs).
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


def grid(n):
    x = np.arange(n) * 2 * np.pi / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    return xx, y, z


def abc_u(x, y, z):
    return np.stack([np.sin(z) + np.cos(y), np.sin(x) + np.cos(z), np.sin(y) + np.cos(x)], axis=-1)


def abc_spectral(n):
    x, y, z = grid(n)
    return np.fft.rfftn(abc_u(x, y, z), axes=(0, 1, 2)).transpose(1, 0, 2, 3) / n ** 3


def make(B, n, precision, dealias, w, nu=0.03, forcing="linear", fk0=2.0, fk1=4.0):
    fs = B.vorticity_equation("s", n, precision=precision, dealias=dealias)
    fs.nu, fs.forcing_type, fs.famplitude, fs.fk0, fs.fk1 = nu, forcing, 0.5, fk0, fk1
    fs.io_checkpoint(read=True, data=w.astype(fs.cdtype))
    return fs


def real_of(B, fs, nc, spec):
    f = B.field(fs, nc)
    f.put(spec, real_space=False)
    f.ift()
    return f.get()


@pytest.mark.parametrize("precision,dealias", [("double", "SMOOTH"), ("single", "SMOOTH"), ("double", "TWO_THIRDS")])
def test_abc_pressure_and_acceleration(B, precision, dealias):
    n, nu = 32, 0.03
    say("case: ABC pressure/acceleration n=%d %s %s" % (n, precision, dealias))
    fs = make(B, n, precision, dealias, abc_spectral(n), nu=nu)
    x, y, z = grid(n)
    p = B.field(fs, "ONE")
    fs.compute_pressure(p)
    p.ift()
    pr = p.get()[..., 0]
    pref = -(np.cos(y) * np.sin(z) + np.cos(z) * np.sin(x) + np.cos(x) * np.sin(y))
    tol = 1e-12 if precision == "double" else 2e-5
    say("  pressure rel", rel(pr, pref))
    assert rel(pr, pref) < tol
    a = B.field(fs, "THREE")
    fs.compute_Lagrangian_acceleration(a)
    a.ift()
    ar = a.get()
    g = np.stack([np.cos(z) * np.cos(x) - np.sin(x) * np.sin(y), np.cos(x) * np.cos(y) - np.sin(y) * np.sin(z),
                  np.cos(y) * np.cos(z) - np.sin(z) * np.sin(x)], axis=-1)
    say("  acceleration rel", rel(ar, g - nu * abc_u(x, y, z)))
    assert rel(ar, g - nu * abc_u(x, y, z)) < tol
    fs.close()


@pytest.mark.parametrize("n,precision,dealias,forcing", [(32, "double", "SMOOTH", "linear"),
                                                         (48, "double", "TWO_THIRDS", "linear"),
                                                         (32, "single", "SMOOTH", "linear"),
                                                         (64, "double", "SMOOTH", "Kolmogorov")])
def test_pressure_and_acceleration_vs_oracle(B, oracle, n, precision, dealias, forcing):
    from bfps_amd import ic
    say("case: pressure/acceleration vs oracle n=%d %s %s %s" % (n, precision, dealias, forcing))
    w = ic.solenoidal_vorticity(n, seed=7, k0=3.0)
    o = oracle.Solver(n, precision="double", dealias=dealias)
    o.set_physics(0.02, forcing, 0.5, 1.0, 3.0, 1)
    o.set_cfield("vorticity", w)
    o.io_projection()
    fs = make(B, n, precision, dealias, w, nu=0.02, forcing=forcing, fk0=1.0, fk1=3.0)
    p = B.field(fs, "ONE")
    fs.compute_pressure(p)
    pref = drop_nyquist(o.pressure()[..., None])
    a = B.field(fs, "THREE")
    fs.compute_Lagrangian_acceleration(a)
    aref = drop_nyquist(o.lagrangian_acceleration())
    tol = 1e-12 if precision == "double" else 2e-5
    ep, ea = rel(p.get(), pref), rel(a.get(), aref)
    say("  pressure rel %.3g acceleration rel %.3g" % (ep, ea))
    assert ep < tol and ea < tol
    fs.close()


@pytest.mark.parametrize("n,nc,precision", [(32, 3, "double"), (48, 9, "single"), (32, 1, "single"),
                                            (64, 3, "single")])
def test_rspace_stats_vs_oracle(B, oracle, n, nc, precision):
    say("case: rspace stats n=%d nc=%d %s" % (n, nc, precision))
    rng = np.random.default_rng(n * 10 + nc)
    fs = B.vorticity_equation("s", n, precision=precision)
    r = rng.standard_normal((n, n, n, nc)).astype(fs.rdtype)
    f = B.field(fs, nc)
    f.put(r, real_space=True)
    nvals = 4 if nc == 3 else nc
    me = np.full(nvals, 2.5)
    if nc == 3:
        me[3] = 2.5 * np.sqrt(3.0)
    for nbins in (64, 1000):
        mom, hist = f.compute_rspace_stats(me, nbins)
        rmom, rhist = oracle.rspace_stats(r.astype(np.float64), nbins, me)
        assert np.array_equal(hist, rhist), np.abs(hist - rhist).max()
        assert np.array_equal(mom[[0, 9]], rmom[[0, 9]])
        assert np.allclose(mom[1:9], rmom[1:9], rtol=1e-12, atol=1e-15)
        say("  nbins %d: histograms equal, moments rel %.3g" % (nbins, rel(mom, rmom)))
    fs.close()


def test_rspace_stats_closed_forms(B):
    n = 16
    x, y, z = grid(n)
    fs = B.vorticity_equation("s", n, precision="double")
    f = B.field(fs, 1)
    f.put(np.sin(x)[..., None], real_space=True)
    mom, hist = f.compute_rspace_stats([1.0], 8)
    assert mom[0, 0] == -1.0 and mom[9, 0] == 1.0
    assert abs(mom[2, 0] - 0.5) < 1e-15 and abs(mom[4, 0] - 0.375) < 1e-15
    assert hist.sum() == n ** 3 - n * n  # sin = 1 falls in bin nbins (dropped)
    f.put(np.full((n, n, n, 1), -0.25), real_space=True)
    mom, hist = f.compute_rspace_stats([1.0], 16)
    assert mom[0, 0] == -0.25 and mom[9, 0] == 0.0  # max starts at 0
    assert hist[int((-0.25 + 1.0) / (2.0 / 16)), 0] == n ** 3
    fs.close()


def test_compute_stats_vs_oracle(B, oracle):
    from bfps_amd import ic
    n = 32
    say("case: compute_stats n=%d" % n)
    w = ic.solenoidal_vorticity(n, seed=3, k0=2.0)
    fs = make(B, n, "double", "SMOOTH", w)
    o = oracle.Solver(n, precision="double")
    o.set_cfield("vorticity", w)
    o.io_projection()
    o.velocity_rspace()
    ur = o.get_rfield("rvelocity")
    # spectral input (NSVE::do_stats): cospectrum, then ift and real stats
    f = B.field(fs, "THREE")
    f.copy_from("velocity")
    st = f.compute_stats(0.5, 128)
    assert f.real_space_representation
    o.compute_velocity()
    assert rel(st["cospectrum"], o.cospectrum("velocity", "velocity")) < 1e-12
    me = [0.5] * 3 + [0.5 * np.sqrt(3.0)]
    rmom, rhist = oracle.rspace_stats(ur, 128, me)
    # the real field is the GPU's own transform: histograms may differ where
    # a value sits within rounding of a bin edge
    assert np.abs(st["histograms"] - rhist).sum() <= 4
    assert np.allclose(st["moments"], rmom, rtol=1e-10, atol=1e-14)
    # real input: real stats, then dft + normalize, cospectrum
    g = B.field(fs, "THREE")
    g.put(ur, real_space=True)
    st2 = g.compute_stats(0.5, 128)
    assert not g.real_space_representation
    assert np.array_equal(st2["histograms"], rhist)
    assert rel(st2["cospectrum"], o.cospectrum("velocity", "velocity")) < 1e-12
    fs.close()


def test_stats_errors(B):
    fs = B.vorticity_equation("s", 16, precision="double")
    f = B.field(fs, 3)
    with pytest.raises(B.capi.BfpshError):
        f.compute_rspace_stats([1, 1, 1, 1], 16)  # spectral representation
    f.ift()
    with pytest.raises(B.capi.BfpshError):
        f.compute_rspace_stats([1, 1, 1, 1], 0)
    with pytest.raises(B.capi.BfpshError):
        f.compute_rspace_stats([1, 1, 0, 1], 16)
    with pytest.raises(B.capi.BfpshError):
        fs.compute_pressure(f)  # needs one component
    p = B.field(fs, 1)
    with pytest.raises(B.capi.BfpshError):
        fs.compute_Lagrangian_acceleration(p)  # needs three
    fs.close()


def test_nsveparticles_do_stats_abc(B):
    # NSVEparticles::do_stats on the ABC flow: the sampled velocity and
    # Lagrangian acceleration at random tracers against the analytic fields
    # (interpolation n4 m1 on 32^3), the statistics against their closed
    # forms: |u|^2 mean 3, u_x^2 mean 1, vorticity = velocity
    n, nu = 32, 0.03
    fs = make(B, n, "double", "SMOOTH", abc_spectral(n), nu=nu)
    rng = np.random.default_rng(5)
    pos = rng.random((3000, 3)) * 2 * np.pi
    ps = B.particles_system(fs, 3000, 4, 1, 4, step_idx=1)
    ps.put(pos)
    # estimates 3 sqrt(3): component ranges +-3 hold the ABC values, whose
    # maximum 2 would sit on the top edge (dropped) with a range of +-2
    st = B.NSVEparticles_do_stats(fs, ps, 3 * np.sqrt(3.0), 3 * np.sqrt(3.0), 64)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    u = abc_u(x, y, z)
    a = np.stack([np.cos(z) * np.cos(x) - np.sin(x) * np.sin(y), np.cos(x) * np.cos(y) - np.sin(y) * np.sin(z),
                  np.cos(y) * np.cos(z) - np.sin(z) * np.sin(x)], axis=-1) - nu * u
    eu = np.abs(st["tracers0/velocity"] - u).max()
    ea = np.abs(st["tracers0/acceleration"] - a).max()
    say("  sampled velocity err %.3g acceleration err %.3g" % (eu, ea))
    assert eu < 1e-4 and ea < 1e-4
    for name in ("velocity", "vorticity"):
        mom = st[name]["moments"]
        assert abs(mom[2, 3] - 3.0) < 1e-12 and np.allclose(mom[2, :3], 1.0, atol=1e-12)
        assert st[name]["histograms"][:, :3].sum() == 3 * n ** 3
    fs.close()
