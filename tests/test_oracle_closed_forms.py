"""Pin the oracle's solver and tracer functions against what the mathematics
fixes, independently of the oracle's own formulas:

* Beltrami (ABC) flow, omega = u on |k| = 1: u x omega = 0, so one step is
  pure viscous decay exp(-nu dt); its velocity and gradient are analytic
  (bfps/tests/test_particles.py:38-68 builds the same field).
* A single Fourier wave k: u x omega is parallel to k and a function of k.x,
  hence curl-free; with linear forcing on its shell the IF-SSP-RK3 step is
  exactly exp(-nu k^2 dt) (1 + a + a^2/2 + a^3/6), a = A dt (the RK3
  stability polynomial; vorticity_equation.cpp:358-439, add_forcing 241-288).
* Kolmogorov forcing from rest: one step equals Simpson's rule of the
  variation-of-constants integral.
* TWO_THIRDS / SMOOTH projection (kspace.cpp:83-92,235-253,487-519): modes
  with k^2 >= kM^2 vanish, k . omega = 0.
* cospectrum (kspace.cpp:521-598): Parseval against the real-space field,
  and shell binning of a single wave.
* Tracers: interpolation of the analytic ABC velocity converges with the
  stencil size; Adams-Bashforth of order k (particles_adams_bashforth.hpp)
  converges as dt^k in a frozen field against an RK4 trajectory.
"""
import numpy as np
import pytest

TWOPI = 2 * np.pi


def grid(n):
    x = np.arange(n) * TWOPI / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    return xx, y, z


def abc_u(x, y, z):
    """u = (sin z + cos y, sin x + cos z, sin y + cos x); curl u = u."""
    return np.stack([np.sin(z) + np.cos(y), np.sin(x) + np.cos(z), np.sin(y) + np.cos(x)], axis=-1)


def abc_grad(x, y, z):
    """[j][i] = d_j u_i, flattened j*3+i."""
    zero = np.zeros_like(x)
    g = np.zeros(x.shape + (3, 3))
    g[..., 0, 1] = np.cos(x)    # d_x u_y
    g[..., 0, 2] = -np.sin(x)   # d_x u_z
    g[..., 1, 0] = -np.sin(y)   # d_y u_x
    g[..., 1, 2] = np.cos(y)    # d_y u_z
    g[..., 2, 0] = np.cos(z)    # d_z u_x
    g[..., 2, 1] = -np.sin(z)   # d_z u_y
    del zero
    return g.reshape(x.shape + (9,))


def spectral(real):
    """real [z][y][x][c] -> true Fourier coefficients [ky][kz][kx][c]."""
    n = real.shape[0]
    return np.fft.rfftn(real, axes=(0, 1, 2)).transpose(1, 0, 2, 3) / n ** 3


def abc_vorticity(n):
    x, y, z = grid(n)
    return spectral(abc_u(x, y, z))


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("precision,dealias,forcing", [("double", "SMOOTH", "none"), ("double", "TWO_THIRDS", "linear"),
                                                       ("single", "SMOOTH", "linear")])
def test_beltrami_decay_and_analytic_velocity(oracle, precision, dealias, forcing):
    n, nu, dt = 16, 0.05, 0.01
    w0 = abc_vorticity(n)
    s = oracle.Solver(n, precision=precision, dealias=dealias)
    s.set_physics(nu, forcing, 0.5, 2.0, 4.0)   # band 2..4 misses |k| = 1
    s.set_cfield("vorticity", w0)
    s.io_projection()
    x, y, z = grid(n)
    s.velocity_rspace()
    tol = 1e-13 if precision == "double" else 2e-6
    assert rel(s.get_rfield("rvelocity"), abc_u(x, y, z)) < tol
    s.gradient_rspace()
    assert rel(s.get_rfield("gradient"), abc_grad(x, y, z)) < tol
    for _ in range(3):
        s.step(dt)
    assert rel(s.get_cfield("vorticity"), w0 * np.exp(-nu * 3 * dt)) < tol


def single_wave(n, k, amp):
    """Vorticity made of the single wave vector k = (kx, ky, kz), kx > 0, with
    complex polarisation amp (made perpendicular to k)."""
    kv = np.array(k, dtype=float)
    a = np.asarray(amp, dtype=complex)
    a = a - kv * (kv @ a) / (kv @ kv)
    w = np.zeros((n, n, n // 2 + 1, 3), dtype=complex)
    kx, ky, kz = k
    w[ky % n, kz % n, kx] = a
    return w


@pytest.mark.parametrize("k,inband", [((1, 2, 0), True), ((2, -1, 3), True), ((1, 0, 0), False), ((3, 3, 3), False)])
def test_single_wave_linear_forcing_rk3_polynomial(oracle, k, inband):
    n, nu, dt, A = 16, 0.02, 0.05, 0.5
    w0 = single_wave(n, k, [0.3 + 0.1j, -0.2 + 0.4j, 0.25 - 0.3j])
    s = oracle.Solver(n, precision="double")
    s.set_physics(nu, "linear", A, 2.0, 4.0)
    s.set_cfield("vorticity", w0)
    s.io_projection()
    s.step(dt)
    k2 = float(np.dot(k, k))
    a = A * dt if inband else 0.0
    expect = w0 * np.exp(-nu * k2 * dt) * (1 + a + a * a / 2 + a ** 3 / 6)
    got = s.get_cfield("vorticity")
    assert rel(got, expect) < 1e-13
    # a plausible slip (forcing left out, or Euler instead of RK3) is visible
    if inband:
        assert rel(w0 * np.exp(-nu * k2 * dt) * (1 + a), expect) > 1e-4


def test_kolmogorov_from_rest_is_simpson(oracle):
    n, nu, dt, A, fm = 16, 0.1, 0.05, 0.7, 2
    s = oracle.Solver(n, precision="double")
    s.set_physics(nu, "Kolmogorov", A, 2.0, 4.0, fm)
    s.set_cfield("vorticity", np.zeros((n, n, n // 2 + 1, 3), dtype=complex))
    s.io_projection()
    lam = nu * fm * fm * dt
    S = (1 + np.exp(-lam)) / 6 + 2 * np.exp(-lam / 2) / 3
    expect = np.zeros((n, n, n // 2 + 1, 3), dtype=complex)
    v = 0.0
    for _ in range(3):
        s.step(dt)
        v = v * np.exp(-lam) - A / 2 * dt * S
        # forcing on the z component of the (ky = +-fmode, kz = 0, kx = 0) modes
        expect[fm, 0, 0, 2] = v
        expect[n - fm, 0, 0, 2] = v
        assert np.abs(s.get_cfield("vorticity") - expect).max() < 1e-15


@pytest.mark.parametrize("n,dealias,kM", [(48, "TWO_THIRDS", 15.0), (32, "SMOOTH", 15.0), (24, "TWO_THIRDS", 7.0)])
def test_projection_lowpass_and_divfree(oracle, n, dealias, kM):
    rng = np.random.default_rng(3)
    shp = (n, n, n // 2 + 1, 3)
    w = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    s = oracle.Solver(n, precision="double", dealias=dealias)
    assert s.kM == kM and s.kM2 == kM * kM
    assert s.nshells == int(kM) + 2
    s.set_cfield("vorticity", w)
    s.io_projection()
    g = s.get_cfield("vorticity")
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    KY, KZ, KX = np.meshgrid(k1, k1, kx, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    assert np.abs(g[K2 >= kM * kM]).max() == 0
    div = KX * g[..., 0] + KY * g[..., 1] + KZ * g[..., 2]
    assert np.abs(div).max() < 1e-13 * np.abs(g).max()
    # away from the symmetrised kx = 0 plane and the k = 0 mode, the kept modes
    # are the divergence-free part of the input
    keep = (K2 < kM * kM) & (KX > 0)
    kk = np.stack([KX, KY, KZ], -1)[keep].astype(float)
    wi = w[keep]
    proj = wi - kk * (np.sum(kk * wi, -1) / np.sum(kk * kk, -1))[:, None]
    assert np.abs(g[keep] - proj).max() < 1e-13


def test_cospectrum_parseval_and_binning(oracle):
    n = 32
    rng = np.random.default_rng(5)
    shp = (n, n, n // 2 + 1, 3)
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    K = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kx[None, None, :] ** 2)
    w = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * np.exp(-(K / 4) ** 2)[..., None]
    s = oracle.Solver(n, precision="double")
    s.set_cfield("vorticity", w)
    s.io_projection()
    s.compute_velocity()
    s.velocity_rspace()
    u = s.get_rfield("rvelocity")
    S = s.cospectrum("velocity", "velocity")
    tr = np.trace(S, axis1=1, axis2=2).sum()
    assert abs(tr - np.mean(np.sum(u * u, -1))) < 1e-12 * tr
    assert np.allclose(S, S.transpose(0, 2, 1))
    # all of a single wave's energy lands in shell int(|k| / dk)
    s2 = oracle.Solver(n, precision="double")
    s2.set_cfield("vorticity", single_wave(n, (2, 2, 1), [1.0, 0.5j, -0.3]))
    s2.io_projection()
    s2.compute_velocity()
    S2 = s2.cospectrum("velocity", "velocity")
    tr2 = np.trace(S2, axis1=1, axis2=2)
    assert np.count_nonzero(tr2) == 1 and tr2[3] > 0


def test_kshell_counts(oracle):
    """nshell counts modes with k^2 < kM^2 per shell, kx > 0 modes twice
    (their conjugates); the total is the number of lattice points in the
    open ball of radius kM."""
    n = 32
    s = oracle.Solver(n, precision="double")
    ks, ns = s.kshell()
    k = np.arange(-n // 2, n // 2)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    inside = K2 < s.kM2
    assert ns.sum() == inside.sum()
    shell = (np.sqrt(K2[inside]) / s.dk).astype(int)
    assert np.array_equal(np.bincount(shell, minlength=s.nshells), ns)
    mean_k = np.bincount(shell, weights=np.sqrt(K2[inside]), minlength=s.nshells)
    nz = ns > 0
    assert np.allclose(ks[nz], mean_k[nz] / ns[nz])


def _abc_solver(oracle, n):
    s = oracle.Solver(n, precision="double")
    s.set_physics(0.0, "none")
    s.set_cfield("vorticity", abc_vorticity(n))
    s.io_projection()
    s.velocity_rspace()
    s.gradient_rspace()
    return s


@pytest.mark.parametrize("n,m", [(1, 0), (1, 1), (2, 1), (3, 2), (4, 1), (6, 2), (8, 2)])
def test_interpolation_converges_to_analytic(oracle, n, m):
    """Sampled ABC velocity and gradient against the analytic field at
    random points, including points outside [0, 2pi)."""
    N = 32
    s = _abc_solver(oracle, N)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-2 * TWOPI, 3 * TWOPI, (500, 3))
    p = oracle.Particles(s, len(pos), n, m, 1)
    p.set_state(pos)
    u = p.sample("velocity")
    g = p.sample("gradient")
    ue = abc_u(pos[:, 0], pos[:, 1], pos[:, 2])
    ge = abc_grad(pos[:, 0], pos[:, 1], pos[:, 2])
    h = TWOPI / N
    # order of the interpolant: min(2n + 1, 2m + 1) + 1 (Lagrange when m = 0)
    order = (2 * n + 2) if m == 0 else min(2 * n + 1, 2 * m + 1) + 1
    bound = 2.0 * h ** order
    eu = np.abs(u - ue).max()
    eg = np.abs(g - ge).max()
    assert eu < bound and eg < bound, (eu, eg, bound)
    # the bound is far below the O(h) error of a stencil or cell-index slip
    assert bound < 0.05 * h


def _rk4_traj(x0, t, steps):
    def f(x):
        return abc_u(x[:, 0], x[:, 1], x[:, 2])
    x = x0.copy()
    h = t / steps
    for _ in range(steps):
        k1 = f(x)
        k2 = f(x + h / 2 * k1)
        k3 = f(x + h / 2 * k2)
        k4 = f(x + h * k3)
        x = x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_adams_bashforth_order(oracle, order):
    """Frozen ABC field, exact history: the error after T = 1 falls as dt^k."""
    N, T = 32, 1.0
    s = _abc_solver(oracle, N)
    rng = np.random.default_rng(11)
    x0 = rng.uniform(0, TWOPI, (20, 3))
    errs = []
    for nsteps_t in (10, 20):
        dt = T / nsteps_t
        # history rhs[i] = u(x(-i dt)) for i >= 1, rhs[0] = 0
        rhs = np.zeros((order, len(x0), 3))
        for i in range(1, order):
            xb = _rk4_traj(x0, -i * dt, 200 * i)
            rhs[i] = abc_u(xb[:, 0], xb[:, 1], xb[:, 2])
        p = oracle.Particles(s, len(x0), 8, 2, order, step_idx=100)
        p.set_state(x0, rhs)
        for _ in range(nsteps_t):
            p.complete_loop(dt)
        xe = _rk4_traj(x0, T, 4000)
        errs.append(np.abs(p.get_state()[0] - xe).max())
    ratio = errs[0] / errs[1]
    assert 2 ** order * 0.7 < ratio < 2 ** order * 1.5, (errs, ratio)


@pytest.mark.parametrize("dealias,k1,k2", [("SMOOTH", (2, 3, 0), (2, 2, 1)), ("SMOOTH", (1, 3, -2), (3, 2, 2)),
                                           ("TWO_THIRDS", (2, 1, 0), (1, 1, 1)),
                                           ("TWO_THIRDS", (2, 1, 1), (2, 1, 0))])
def test_triad_nonlinear_term_and_filter(oracle, dealias, k1, k2):
    """Two waves k1, k2 feed k3 = k1 + k2 at the rate
    filter(k3) * i k3 x (u1 x w2 + u2 x w1) (Fourier coefficients, u = i k x w / k^2),
    with filter = exp(-36 (k^2/kM^2)^18) for SMOOTH (kspace.cpp:88-92,135,461-485)
    or the sharp cut k^2 < kM^2 for TWO_THIRDS.  The initial rate is read from
    two short steps by Richardson extrapolation."""
    n = 16
    rng = np.random.default_rng(sum(k1) + 7 * sum(k2))
    w = single_wave(n, k1, rng.standard_normal(3) + 1j * rng.standard_normal(3))
    w += single_wave(n, k2, rng.standard_normal(3) + 1j * rng.standard_normal(3))
    k1v, k2v = np.array(k1, float), np.array(k2, float)
    k3 = k1v + k2v
    w1 = w[k1[1] % n, k1[2] % n, k1[0]]
    w2 = w[k2[1] % n, k2[2] % n, k2[0]]
    u1 = 1j * np.cross(k1v, w1) / (k1v @ k1v)
    u2 = 1j * np.cross(k2v, w2) / (k2v @ k2v)
    s0 = oracle.Solver(n, precision="double", dealias=dealias)
    kM2 = s0.kM2
    k32 = k3 @ k3
    if dealias == "SMOOTH":
        filt = np.exp(-36.0 * (k32 / kM2) ** 18)
    else:
        filt = 1.0 if k32 < kM2 else 0.0
    rate = 1j * np.cross(k3, filt * (np.cross(u1, w2) + np.cross(u2, w1)))
    est = []
    for dt in (1e-4, 2e-4):
        s = oracle.Solver(n, precision="double", dealias=dealias)
        s.set_physics(0.0, "none")
        s.set_cfield("vorticity", w)
        s.step(dt)
        est.append(s.get_cfield("vorticity")[int(k3[1]) % n, int(k3[2]) % n, int(k3[0])] / dt)
    got = 2 * est[0] - est[1]
    assert np.abs(got - rate).max() < 1e-6 * max(np.abs(rate).max(), 1e-3)
    if dealias == "SMOOTH":
        assert 0.05 < filt < 0.95   # the filter value matters for these triads


def band_energy(w, k):
    """1/2 sum nxmodes |omega|^2 / k^2 of a single stored wave."""
    kv = np.array(k, float)
    a = w[k[1] % w.shape[0], k[2] % w.shape[1], k[0]]
    return 0.5 * (2 if k[0] > 0 else 1) * np.sum(np.abs(a) ** 2) / (kv @ kv)


@pytest.mark.parametrize("nu", [0.0, 0.05])
def test_fixed_energy_injection_single_wave(oracle, nu):
    """Forcing with a fixed injection rate eps on the band: for a single wave
    (no nonlinear transfer) dE/dt = eps - 2 nu k^2 E exactly, so
    E(t) = eps/(2 nu k^2) + (E0 - eps/(2 nu k^2)) exp(-2 nu k^2 t)
    (E0 + eps t when nu = 0); IF-RK3 follows it to O(dt^3)."""
    n, eps, T = 16, 0.3, 0.2
    k = (1, 1, 2)   # |k| = 2.449 inside 1.5 <= |k| <= 3.5
    w0 = single_wave(n, k, [0.2 + 0.1j, -0.3j, 0.1])
    k2 = float(np.dot(k, k))
    errs = []
    for nsteps in (10, 20):
        dt = T / nsteps
        s = oracle.Solver(n, precision="double")
        s.set_physics(nu, "fixed_energy_injection_rate", eps, 1.5, 3.5)
        s.set_cfield("vorticity", w0)
        s.io_projection()
        e0 = band_energy(s.get_cfield("vorticity"), k)
        for _ in range(nsteps):
            s.step(dt)
        if nu == 0:
            expect = e0 + eps * T
        else:
            einf = eps / (2 * nu * k2)
            expect = einf + (e0 - einf) * np.exp(-2 * nu * k2 * T)
        w = s.get_cfield("vorticity")
        errs.append(abs(band_energy(w, k) - expect) / expect)
        # the wave keeps its shape: only its amplitude changes
        a0 = w0[k[1], k[2], k[0]]
        a1 = w[k[1], k[2], k[0]]
        assert np.abs(a1 / np.linalg.norm(a1) - a0 / np.linalg.norm(a0)).max() < 1e-12
    # close to the exact energy, and converging at third order in dt
    assert errs[0] < 1e-4, errs
    assert 8 * 0.7 < errs[0] / errs[1] < 8 * 1.4, errs


# ---------------------------------------------------------------- statistics
# compute_pressure / compute_Lagrangian_acceleration
# (vorticity_equation.cpp:441-578) on the ABC flow: u x omega = 0, so
# (u . grad) u = grad |u|^2 / 2 and the pressure Poisson equation gives
# p = -|u|^2 / 2 + const, i.e. p = -(cos y sin z + cos z sin x + cos x sin y)
# with the mean removed; the acceleration is -grad p + nu lap u =
# grad |u|^2 / 2 - nu u (the linear forcing band 2..4 misses |k| = 1).
def abc_pressure(x, y, z):
    return -(np.cos(y) * np.sin(z) + np.cos(z) * np.sin(x) + np.cos(x) * np.sin(y))


def abc_acceleration(x, y, z, nu):
    g = np.stack([np.cos(z) * np.cos(x) - np.sin(x) * np.sin(y),
                  np.cos(x) * np.cos(y) - np.sin(y) * np.sin(z),
                  np.cos(y) * np.cos(z) - np.sin(z) * np.sin(x)], axis=-1)
    return g - nu * abc_u(x, y, z)


def real_of(spec, n):
    """true Fourier coefficients [ky][kz][kx][c] -> real [z][y][x][c]."""
    return np.fft.irfftn(spec.transpose(1, 0, 2, 3), s=(n, n, n), axes=(0, 1, 2)) * n ** 3


@pytest.mark.parametrize("precision,dealias", [("double", "SMOOTH"), ("double", "TWO_THIRDS"), ("single", "SMOOTH")])
def test_abc_pressure_and_lagrangian_acceleration(oracle, precision, dealias):
    n, nu = 24, 0.03
    s = oracle.Solver(n, precision=precision, dealias=dealias)
    s.set_physics(nu, "linear", 0.5, 2.0, 4.0)
    s.set_cfield("vorticity", abc_vorticity(n))
    x, y, z = grid(n)
    tol = 1e-13 if precision == "double" else 3e-6
    p = real_of(s.pressure()[..., None], n)[..., 0]
    assert rel(p, abc_pressure(x, y, z)) < tol
    a = real_of(s.lagrangian_acceleration(), n)
    assert rel(a, abc_acceleration(x, y, z, nu)) < tol
    # the forcing term enters only for the linear type and only in its band:
    # with the band moved onto |k| = 1 the acceleration gains famplitude u
    s.set_physics(nu, "linear", 0.5, 0.5, 1.5)
    a2 = real_of(s.lagrangian_acceleration(), n)
    assert rel(a2, abc_acceleration(x, y, z, nu) + 0.5 * abc_u(x, y, z)) < tol
    s.set_physics(nu, "Kolmogorov", 0.5, 0.5, 1.5)
    a3 = real_of(s.lagrangian_acceleration(), n)
    assert rel(a3, abc_acceleration(x, y, z, nu)) < tol


# field::compute_rspace_stats (field.cpp:661-850): moments [10][nvals] with
# min (row 0) and max (row 9) starting from 0 (the magnitude's min from its
# estimate), rows 1..8 = mean value^n; histograms of width 2 M / nbins from
# -M (the magnitude's from 0 with half that width).
def test_rspace_stats_constant_field(oracle):
    n, c = 8, -0.25
    d = np.full((n, n, n, 1), c)
    mom, hist = oracle.rspace_stats(d, 16, [1.0])
    assert mom[0, 0] == c and mom[9, 0] == 0.0          # max starts at 0
    assert np.allclose(mom[1:9, 0], [c ** k for k in range(1, 9)], rtol=1e-15, atol=0)
    want = np.zeros((16, 1), dtype=np.int64)
    want[int(np.floor((c + 1.0) / (2.0 / 16))), 0] = n ** 3
    assert np.array_equal(hist, want)


def test_rspace_stats_sine(oracle):
    # u_x = sin x on an 8-point period: the values are 0, s, 1, s, 0, -s,
    # -1, -s (s = 1/sqrt 2); mean sin^k over a full period is
    # C(k, k/2) / 2^k for even k < 8 and 0 for odd k
    n = 8
    x, y, z = grid(n)
    d = np.sin(x)[..., None]
    mom, hist = oracle.rspace_stats(d, 4, [1.0])
    from math import comb
    want = [0.0 if k % 2 else comb(k, k // 2) / 2 ** k for k in range(1, 8)]
    assert np.allclose(mom[1:8, 0], want, atol=1e-15)
    assert mom[0, 0] == -1.0 and mom[9, 0] == 1.0
    # bins of width 0.5 from -1: {-1, -s, -s} -> 0, {0, 0} -> 2, {s, s} -> 3,
    # 1 -> bin 4 = nbins (dropped); per x-period, times n^2 (y, z)
    assert list(hist[:, 0]) == [3 * n * n, 0, 2 * n * n, 2 * n * n]


def test_rspace_stats_unit_magnitude(oracle):
    # u = (cos z, sin z, 0): |u| = 1 everywhere
    n = 8
    x, y, z = grid(n)
    d = np.stack([np.cos(z), np.sin(z), np.zeros_like(z)], axis=-1)
    me = [1.0, 1.0, 1.0, np.sqrt(3.0)]
    mom, hist = oracle.rspace_stats(d, 32, me)
    assert mom.shape == (10, 4) and hist.shape == (32, 4)
    assert np.allclose(mom[:, 3], 1.0, atol=1e-15)       # min |u| = min(sqrt 3, 1)
    assert np.allclose(mom[2, :3], [0.5, 0.5, 0.0], atol=1e-15)
    assert mom[0, 2] == 0.0 and mom[9, 2] == 0.0
    b = int(np.floor(1.0 * 32 / np.sqrt(3.0)))            # magnitude bins of width sqrt 3 / 32
    assert hist[b, 3] == n ** 3 and hist[:, 3].sum() == n ** 3
    assert hist[16, 2] == n ** 3                          # u_z = 0 -> bin 16 of 32 on [-1, 1)


# ---------------------------------------------------------------- compute_vorticity
def test_compute_vorticity_closed_forms(oracle):
    """compute_vorticity (vorticity_equation.cpp:171-204) against curls fixed
    by the mathematics: the ABC flow is its own curl; u = (sin 2y cos z, 0,
    cos x) has curl (0, sin x - sin 2y sin z, -2 cos 2y cos z) (transformed
    by numpy, not by the oracle)."""
    n = 16
    o = oracle.Solver(n, precision="double")
    u = abc_vorticity(n)
    o.set_cfield("velocity", u)
    o.compute_vorticity()
    assert rel(o.get_cfield("vorticity"), u) < 1e-14
    x, y, z = grid(n)
    ur = np.stack([np.sin(2 * y) * np.cos(z), np.zeros_like(x), np.cos(x)], axis=-1)
    wr = np.stack([np.zeros_like(x), np.sin(x) - np.sin(2 * y) * np.sin(z), -2 * np.cos(2 * y) * np.cos(z)], axis=-1)
    o.set_cfield("velocity", spectral(ur))
    o.compute_vorticity()
    assert rel(o.get_cfield("vorticity"), spectral(wr)) < 1e-14


@pytest.mark.parametrize("dealias", ["SMOOTH", "TWO_THIRDS"])
def test_compute_vorticity_inverts_biot_savart(oracle, dealias):
    """compute_velocity(compute_vorticity(u)) = u for a solenoidal u with a
    zero mean and modes only below kM; modes at k^2 > kM^2 of the curl of a
    broadband u are zero."""
    n = 24
    o = oracle.Solver(n, precision="double", dealias=dealias)
    rng = np.random.default_rng(4)
    shp = (n, n, n // 2 + 1, 3)
    u = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    K = np.stack(np.meshgrid(k1, k1, kx, indexing="ij"), -1)[..., [2, 0, 1]].astype(float)  # (kx, ky, kz)
    k2 = (K ** 2).sum(-1)
    # broadband input: the curl vanishes beyond kM
    o.set_cfield("velocity", u)
    o.compute_vorticity()
    w = o.get_cfield("vorticity")
    assert np.abs(w[k2 > o.kM2]).max() == 0.0
    # band-limited solenoidal input (projected here, and made Hermitian by
    # the field's own symmetrize): the Biot-Savart of the curl returns it
    u = u * (k2 < o.kM2 * 0.8)[..., None]
    u -= K * ((K * u).sum(-1) / np.where(k2 > 0, k2, 1))[..., None]
    u[0, 0, 0] = 0
    u = oracle.symmetrize(u, n)
    o.set_cfield("velocity", u)
    o.compute_vorticity()
    w = o.get_cfield("vorticity")
    o.set_cfield("vorticity", w)
    o.compute_velocity()
    assert rel(o.get_cfield("velocity"), u) < 1e-14
