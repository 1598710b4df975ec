"""GPU parity tests: the HIP backend (through the C ABI) against the CPU oracle,
the reference's golden checkpoint, and closed forms.

Tolerances:
  * FP64 fields: relative L2 <= 1e-11 against the FP64 oracle (rounding only).
  * FP32 fields: relative L2 <= 2e-6 against the FP32-field oracle after a
    few steps; the reference's own 1e-5 max-abs criterion on its fixture
    (bfps/test/test_bfps_NSVEparticles.py:41-50).
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


def random_vorticity(n, seed, k0=3.0, dtype=np.complex128):
    """Band-limited random spectral vorticity in the reference layout."""
    rng = np.random.default_rng(seed)
    shp = (n, n, n // 2 + 1, 3)
    a = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    kx = np.arange(n // 2 + 1)
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    K = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kx[None, None, :] ** 2)
    a *= (np.exp(-(K / k0) ** 2) * (K > 0))[..., None]
    return a.astype(dtype)


NU = (2.0 * 2 / 32) ** (4. / 3)
DT = 0.5 / 32


def test_fixture_nsveparticles_two_jobs(B, oracle, golden):
    """The reference's own regression (32^3 single, 10^4 tracers, n1 m1, AB4,
    2 x 32 iterations with a restart), asserted at the reference's 1e-5."""
    say("case: golden fixture 32^3 single, 64 steps, restart at 32")
    fs = B.vorticity_equation("fixture", 32, precision="single")
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = NU, 0.5, 2.0, 4.0, "linear"
    fs.io_checkpoint(read=True, data=golden["vort_0"])
    ps = B.particles_system(fs, 10000, 1, 1, 4, step_idx=1)
    ps.put(golden["state_0"], golden["rhs_0"])
    for _ in range(32):
        B.NSVEparticles_step(fs, DT)
    v = fs.cvorticity
    pos, rhs, idx, si = ps.get()
    say("it32 errors", np.abs(v - golden["vort_32"]).max(), np.abs(pos - golden["state_32"]).max(),
        np.abs(rhs - golden["rhs_32"]).max())
    assert si == 33
    assert np.abs(v - golden["vort_32"]).max() < 1e-5
    assert np.abs(pos - golden["state_32"]).max() < 1e-5
    assert np.abs(rhs - golden["rhs_32"]).max() < 1e-5
    fs2 = B.vorticity_equation("fixture2", 32, precision="single")
    fs2.nu, fs2.famplitude, fs2.fk0, fs2.fk1, fs2.forcing_type = NU, 0.5, 2.0, 4.0, "linear"
    fs2.io_checkpoint(read=True, data=v)
    ps2 = B.particles_system(fs2, 10000, 1, 1, 4, step_idx=33)
    ps2.put(pos, rhs)
    for _ in range(32):
        B.NSVEparticles_step(fs2, DT)
    v = fs2.cvorticity
    pos, rhs, idx, si = ps2.get()
    say("it64 errors", np.abs(v - golden["vort_64"]).max(), np.abs(pos - golden["state_64"]).max(),
        np.abs(rhs - golden["rhs_64"]).max())
    assert si == 65
    assert np.abs(v - golden["vort_64"]).max() < 1e-5
    assert np.abs(pos - golden["state_64"]).max() < 1e-5
    assert np.abs(rhs - golden["rhs_64"]).max() < 1e-5


@pytest.mark.parametrize("n", [16, 32, 48, 64])
@pytest.mark.parametrize("dealias", ["SMOOTH", "TWO_THIRDS"])
def test_step_fp64_vs_oracle(B, oracle, n, dealias):
    say("case: fp64 step n=%d dealias=%s" % (n, dealias))
    w0 = random_vorticity(n, seed=n)
    fs = B.vorticity_equation("t", n, precision="double", dealias=dealias)
    fs.nu, fs.famplitude, fs.forcing_type = 0.02, 0.3, "linear"
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="double", dealias=dealias)
    o.set_physics(0.02, "linear", 0.3, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    e0 = rel(fs.cvorticity, o.get_cfield("vorticity"))
    say("  after projection rel", e0)
    assert e0 < 1e-14
    for _ in range(4):
        fs.step(0.01)
        o.step(0.01)
    e = rel(fs.cvorticity, o.get_cfield("vorticity"))
    say("  after 4 steps rel", e)
    assert e < 1e-11


@pytest.mark.parametrize("n", [32, 64, 128])
def test_step_fp32_vs_oracle(B, oracle, n):
    """FP32 fields on a resolved flow (energy 1/2, k0 = 4, dt = 0.5/N): the
    GPU's error against the FP64 oracle must be of the size of the oracle's
    own FP32-vs-FP64 difference (rounding only), and small in absolute terms."""
    from bfps_amd import ic
    say("case: fp32 step n=%d" % n)
    w0 = ic.solenoidal_vorticity(n, seed=7 + n, k0=4.0)
    nu, dt = 0.05 * (32.0 / n), 0.5 / n
    fs = B.vorticity_equation("t", n, precision="single")
    fs.nu, fs.famplitude, fs.forcing_type = nu, 0.5, "linear"
    fs.io_checkpoint(read=True, data=w0)
    o32 = oracle.Solver(n, precision="single")
    o64 = oracle.Solver(n, precision="double")
    for o in (o32, o64):
        o.set_physics(nu, "linear", 0.5, 2.0, 4.0)
        o.set_cfield("vorticity", w0)
        o.io_projection()
    for _ in range(8):
        fs.step(dt)
        o32.step(dt)
        o64.step(dt)
    ref = o64.get_cfield("vorticity")
    e_gpu = rel(fs.cvorticity, ref)
    e_orc = rel(o32.get_cfield("vorticity"), ref)
    e_pair = rel(fs.cvorticity, o32.get_cfield("vorticity"))
    say("  8 steps: gpu32 vs orc64 %.3e, orc32 vs orc64 %.3e, gpu32 vs orc32 %.3e" % (e_gpu, e_orc, e_pair))
    assert e_gpu < 1e-6
    assert e_gpu < 3 * e_orc + 1e-7


@pytest.mark.parametrize("forcing", ["none", "Kolmogorov", "fixed_energy_injection_rate"])
def test_forcing_variants_fp64(B, oracle, forcing):
    n = 32
    say("case: forcing %s" % forcing)
    w0 = random_vorticity(n, seed=3)
    fs = B.vorticity_equation("t", n, precision="double")
    fs.nu, fs.famplitude, fs.forcing_type, fs.fmode = 0.05, 0.7, forcing, 2
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="double")
    o.set_physics(0.05, forcing, 0.7, 2.0, 4.0, 2)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    for _ in range(3):
        fs.step(0.02)
        o.step(0.02)
    e = rel(fs.cvorticity, o.get_cfield("vorticity"))
    say("  rel", e)
    assert e < 1e-11


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-6)])
def test_velocity_and_gradient_rspace(B, oracle, precision, tol):
    n = 48
    say("case: velocity/gradient rspace n=48 %s" % precision)
    w0 = random_vorticity(n, seed=11)
    if precision == "single":
        w0 = w0.astype(np.complex64)
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision=precision)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    assert rel(fs.compute_velocity(), (o.compute_velocity(), o.get_cfield("velocity"))[1]) < tol
    fs.velocity_rspace()
    o.velocity_rspace()
    e = rel(fs.get_rfield(B.capi.RVELOCITY), o.get_rfield("rvelocity"))
    say("  velocity rel", e)
    assert e < tol
    fs.gradient_rspace()
    o.gradient_rspace()
    e = rel(fs.get_rfield(B.capi.RGRADIENT), o.get_rfield("gradient"))
    say("  gradient rel", e)
    assert e < tol


def test_cospectrum(B, oracle):
    n = 32
    say("case: cospectrum")
    w0 = random_vorticity(n, seed=5)
    fs = B.vorticity_equation("t", n, precision="double")
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="double")
    o.set_cfield("vorticity", w0)
    o.io_projection()
    o.compute_velocity()
    for a, b in [("velocity", "velocity"), ("vorticity", "vorticity"), ("velocity", "vorticity")]:
        g = fs.cospectrum(a, b)
        r = o.cospectrum(a, b)
        assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()
    ks, ns = o.kshell()
    assert np.allclose(fs.kk.kshell, ks) and np.array_equal(fs.kk.nshell, ns)


# FP64: n <= 6 run the LDS-brick kernel (n <= 4 with the weights stored per
# batch, 5 and 6 evaluated per component), n >= 7 the global-gather kernel;
# n = 1, m = 2 has polynomials of degree 5 over a 4-point stencil
@pytest.mark.parametrize("nb,m", [(1, 0), (1, 1), (1, 2), (2, 2), (4, 1), (6, 2), (3, 0), (5, 1), (7, 1), (8, 2),
                                  (10, 1)])
def test_tracers_vs_oracle(B, oracle, nb, m):
    n = 32
    say("case: tracers n=%d m=%d" % (nb, m))
    w0 = random_vorticity(n, seed=9)
    fs = B.vorticity_equation("t", n, precision="double")
    fs.nu, fs.forcing_type = 0.05, "linear"
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="double")
    o.set_physics(0.05, "linear", 1.0, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    rng = np.random.default_rng(nb * 10 + m)
    npart = 3000
    pos = rng.random((npart, 3)) * 2 * np.pi
    pos[:10] -= 2 * np.pi  # negative coordinates exercise the periodic wrap
    pos[10:20] += 4 * np.pi
    ps = B.particles_system(fs, npart, nb, m, 4, step_idx=1)
    ps.put(pos)
    op = oracle.Particles(o, npart, nb, m, 4, step_idx=1)
    op.set_state(pos)
    for _ in range(5):
        B.NSVEparticles_step(fs, 0.01)
        oracle.nsve_particles_step(o, op, 0.01)
    gp, grhs, gidx, gsi = ps.get()
    op_pos, op_rhs, osi = op.get_state()
    say("  pos err", np.abs(gp - op_pos).max(), "rhs err", np.abs(grhs - op_rhs).max())
    assert gsi == osi == 6
    assert np.abs(gp - op_pos).max() < 1e-12
    assert np.abs(grhs - op_rhs).max() < 1e-12
    # sampling velocity and gradient at the current positions
    fs.velocity_rspace()
    o.velocity_rspace()
    assert np.abs(ps.sample("velocity") - op.sample("velocity")).max() < 1e-12
    fs.gradient_rspace()
    o.gradient_rspace()
    assert np.abs(ps.sample("gradient") - op.sample("gradient")).max() < 1e-11


@pytest.mark.parametrize("n,precision", [(64, "double"), (256, "single")])
def test_beltrami_exact_decay(B, n, precision):
    """ABC flow u = (cos y + sin z, cos z + sin x, cos x + sin y) has
    omega = u, so u x omega = 0 and, without forcing, omega(t) = omega(0)
    exp(-nu t) (|k| = 1 modes; exact integrating factor)."""
    say("case: Beltrami decay n=%d %s" % (n, precision))
    x = np.arange(n) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.cos(Y) + np.sin(Z), np.cos(Z) + np.sin(X), np.cos(X) + np.sin(Y)], axis=-1)
    w = np.fft.rfftn(u, axes=(0, 1, 2)).transpose(1, 0, 2, 3) / n ** 3  # omega = u
    fs = B.vorticity_equation("abc", n, precision=precision)
    fs.nu, fs.forcing_type = 0.1, "none"
    fs.io_checkpoint(read=True, data=w)
    dt, nsteps = 0.01, 10
    for _ in range(nsteps):
        fs.step(dt)
    got = fs.cvorticity
    exact = w * np.exp(-0.1 * dt * nsteps)
    e = rel(got, exact)
    say("  rel err vs exact", e)
    assert e < (1e-13 if precision == "double" else 1e-6)


def test_errors_are_reported(B):
    with pytest.raises(B.capi.BfpshError):
        B.vorticity_equation("bad", 40)
    fs = B.vorticity_equation("t", 16, precision="double")
    with pytest.raises(B.capi.BfpshError):
        B.particles_system(fs, 10, 11, 1)


@pytest.mark.parametrize("n,nb,m", [(64, 6, 2), (64, 4, 1), (48, 1, 1), (48, 1, 2), (64, 5, 2), (48, 7, 1),
                                    (32, 10, 2)])
def test_tracers_fp32_bricks(B, oracle, n, nb, m):
    """FP32 fields, many tracers per brick and several bricks per axis (n <=
    7, weights stored per batch; n >= 8 the global-gather kernel; the
    velocity loop of the 64^3 boxes with n <= 6 is the column walk); the
    interpolated velocities / gradients are FP64 sums of FP32 field values."""
    from bfps_amd import ic
    say("case: fp32 tracers n=%d nb=%d m=%d" % (n, nb, m))
    w0 = ic.solenoidal_vorticity(n, seed=21, k0=4.0)
    fs = B.vorticity_equation("t", n, precision="single")
    fs.nu, fs.forcing_type = 0.01, "linear"
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="single")
    o.set_physics(0.01, "linear", 1.0, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    npart = 20000
    pos = ic.uniform_positions(npart, seed=5)
    ps = B.particles_system(fs, npart, nb, m, 4, step_idx=1)
    ps.put(pos)
    op = oracle.Particles(o, npart, nb, m, 4, step_idx=1)
    op.set_state(pos)
    dt = 0.5 / n
    for _ in range(3):
        B.NSVEparticles_step(fs, dt)
        oracle.nsve_particles_step(o, op, dt)
    gp, grhs, _, gsi = ps.get()
    opos, orhs, osi = op.get_state()
    say("  pos err %.3e rhs err %.3e" % (np.abs(gp - opos).max(), np.abs(grhs - orhs).max()))
    assert gsi == osi
    # FP32 fluid fields differ by rounding between the two FFTs
    assert np.abs(grhs - orhs).max() < 1e-5 * max(1.0, np.abs(orhs).max())
    assert np.abs(gp - opos).max() < 1e-7
    fs.velocity_rspace()
    o.velocity_rspace()
    sv, ov = ps.sample("velocity"), op.sample("velocity")
    assert np.abs(sv - ov).max() < 1e-5 * np.abs(ov).max()
    fs.gradient_rspace()
    o.gradient_rspace()
    sg, og = ps.sample("gradient"), op.sample("gradient")
    say("  gradient sample rel err %.3e" % (np.abs(sg - og).max() / np.abs(og).max()))
    assert np.abs(sg - og).max() < 1e-5 * np.abs(og).max()


def _fp32_tracer_run(B, n, nb, m, pos, kernel, monkeypatch, seed=21, steps=3, precision="single"):
    """A box with tracers advanced `steps` NSVEparticles steps by the
    velocity-loop kernel `kernel` (forced: "column" walk or "brick"s), then
    sampled (velocity, gradient);
    returns (positions, history, step index, velocity, gradient)."""
    from bfps_amd import ic
    monkeypatch.setenv("BFPSH_TRACER_KERNEL", kernel)
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.nu, fs.forcing_type = 0.01, "linear"
    fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=seed, k0=4.0, dtype=fs.cdtype))
    ps = B.particles_system(fs, len(pos), nb, m, 4, step_idx=1)
    ps.put(pos)
    monkeypatch.delenv("BFPSH_TRACER_KERNEL", raising=False)
    dt = 0.5 / n
    fs.profile(True)
    for _ in range(steps):
        B.NSVEparticles_step(fs, dt)
    fs.velocity_rspace()
    sv = ps.sample("velocity")
    fs.gradient_rspace()
    sg = ps.sample("gradient")
    prof = fs.profile_report()
    # the column walk is profiled as *_col, bricks without the suffix
    col = kernel == "column"
    assert ("interp_ab_col" in prof) == col and ("interp_ab" in prof) == (not col)
    assert ("sample_col" in prof) == col and ("sample9_col" in prof) == col
    gp, grhs, _, gsi = ps.get()
    fs.close()
    return gp, grhs, gsi, sv, sg


@pytest.mark.parametrize("n,nb,m", [(128, 6, 2), (128, 4, 1), (128, 1, 1), (128, 2, 2), (128, 3, 0), (128, 5, 2),
                                    (256, 6, 2), (256, 4, 1)])
def test_tracers_fp32_columns(B, oracle, monkeypatch, n, nb, m):
    """FP32 boxes of 64^3 and more advance and sample their tracers by the
    column walk (y-stepping LDS ring of planes; n <= 6, and n <= 4 with
    8-cell steps); against the oracle (as the brick kernel: FP64 sums of
    FP32 field values) and against the brick kernel on the same case
    (trajectories, velocity and gradient samples: the same sums in another
    order, within 1e-12 relative).  Tracers include negative and
    out-of-box coordinates, a dense cluster in one column (several batches
    per y step) and empty columns (the cluster's share)."""
    from bfps_amd import ic
    say("case: fp32 column walk n=%d nb=%d m=%d" % (n, nb, m))
    npart = 30000
    pos = ic.uniform_positions(npart, seed=5 + nb)
    h = 2 * np.pi / n
    rng = np.random.default_rng(nb)
    pos[:20] -= 2 * np.pi
    pos[20:40] += 4 * np.pi
    # 3000 tracers in a 4 x 8 x 4-cell block: up to ~100 per y step
    pos[1000:4000] = np.array([5.5, 40.0, 9.5]) * h + rng.random((3000, 3)) * np.array([4, 8, 4]) * h
    # exactly on cell faces
    pos[4000:4100] = np.round(pos[4000:4100] / h) * h
    gp, grhs, gsi, gv, gg = _fp32_tracer_run(B, n, nb, m, pos, "column", monkeypatch)
    bp, brhs, bsi, bv, bg = _fp32_tracer_run(B, n, nb, m, pos, "brick", monkeypatch)
    scale = max(1.0, np.abs(brhs).max())
    say("  column vs brick: pos %.3e rhs %.3e velocity %.3e gradient %.3e"
        % (np.abs(gp - bp).max(), np.abs(grhs - brhs).max() / scale, np.abs(gv - bv).max() / np.abs(bv).max(),
           np.abs(gg - bg).max() / np.abs(bg).max()))
    assert gsi == bsi == 4
    assert np.abs(grhs - brhs).max() <= 1e-12 * scale
    assert np.abs(gp - bp).max() <= 1e-12 * max(1.0, np.abs(bp).max())
    assert np.abs(gv - bv).max() <= 1e-12 * np.abs(bv).max()
    assert np.abs(gg - bg).max() <= 1e-12 * np.abs(bg).max()
    if n <= 128:
        o = oracle.Solver(n, precision="single")
        o.set_physics(0.01, "linear", 1.0, 2.0, 4.0)
        o.set_cfield("vorticity", ic.solenoidal_vorticity(n, seed=21, k0=4.0))
        o.io_projection()
        op = oracle.Particles(o, npart, nb, m, 4, step_idx=1)
        op.set_state(pos)
        for _ in range(3):
            oracle.nsve_particles_step(o, op, 0.5 / n)
        opos, orhs, osi = op.get_state()
        say("  column vs oracle: pos %.3e rhs %.3e" % (np.abs(gp - opos).max(), np.abs(grhs - orhs).max()))
        assert osi == gsi
        assert np.abs(grhs - orhs).max() < 1e-5 * max(1.0, np.abs(orhs).max())
        assert np.abs(gp - opos).max() < 1e-7


@pytest.mark.parametrize("precision,n,nb,m", [("double", 48, 4, 1), ("single", 48, 4, 1), ("double", 32, 6, 2)])
def test_brick_first_batches_vs_gather(B, monkeypatch, precision, n, nb, m):
    """Regression test for the brick kernel's weight table: every workgroup's
    first batch evaluates its weights from the table copied into LDS, so the
    copy must be complete before that batch's prologue.  Boxes with fewer
    bricks than workgroups make every batch a first batch; 20 steps with the
    field passes (other kernels' LDS contents) in between; the brick kernel's
    trajectories must match the global-gather kernel's."""
    from bfps_amd import ic
    say("case: first batches %s n=%d nb=%d m=%d" % (precision, n, nb, m))
    pos = ic.uniform_positions(30000, seed=77 + nb)
    out = {}
    for kern in ("brick", "gather"):
        monkeypatch.setenv("BFPSH_TRACER_KERNEL", kern)
        fs = B.vorticity_equation("t", n, precision=precision)
        fs.nu, fs.forcing_type = 0.01, "linear"
        fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=3, k0=3.0, dtype=fs.cdtype))
        ps = B.particles_system(fs, len(pos), nb, m, 4, step_idx=1)
        ps.put(pos)
        monkeypatch.delenv("BFPSH_TRACER_KERNEL", raising=False)
        for _ in range(20):
            B.NSVEparticles_step(fs, 0.5 / n)
        out[kern] = ps.get()
        fs.close()
    d = np.abs(out["brick"][0] - out["gather"][0]).max()
    r = np.abs(out["brick"][1] - out["gather"][1]).max() / np.abs(out["gather"][1]).max()
    say("  brick vs gather: pos %.3e rhs %.3e" % (d, r))
    tol = 1e-12 if precision == "double" else 1e-9
    assert d < tol and r < tol


@pytest.mark.parametrize("precision,n,nb,m", [("single", 64, 2, 1), ("double", 64, 4, 1), ("double", 32, 6, 2)])
def test_timed_form_choice_leaves_the_state(B, monkeypatch, precision, n, nb, m):
    """At MEASURE rigor the first tracer call times the velocity loop of
    every applicable kernel form on a copy of the tracer state.  Here that
    first call is a gradient sample (the real velocity buffer is not
    allocated yet, so the trials interpolate the gradient's first three
    planes); then 3 steps and both samples.  The same run with the chosen
    form forced (no trials) must give the same positions, history and
    samples bit for bit: the trials leave the real state as it was."""
    from bfps_amd import ic
    say("case: timed form choice %s n=%d nb=%d m=%d" % (precision, n, nb, m))
    pos = ic.uniform_positions(20000, seed=91 + nb)
    dt = 0.5 / n

    def run(kernel):
        if kernel:
            monkeypatch.setenv("BFPSH_TRACER_KERNEL", kernel)
        fs = B.vorticity_equation("t", n, precision=precision, plan_rigor="FFTW_MEASURE")
        fs.nu, fs.forcing_type = 0.01, "linear"
        fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=5, k0=3.0, dtype=fs.cdtype))
        ps = B.particles_system(fs, len(pos), nb, m, 4, step_idx=1)
        ps.put(pos)
        monkeypatch.delenv("BFPSH_TRACER_KERNEL", raising=False)
        fs.profile(True)
        fs.gradient_rspace()
        g0 = ps.sample("gradient")
        for _ in range(3):
            B.NSVEparticles_step(fs, dt)
        fs.velocity_rspace()
        v = ps.sample("velocity")
        fs.gradient_rspace()
        g = ps.sample("gradient")
        prof = fs.profile_report()
        gp, grhs, _, gsi = ps.get()
        fs.close()
        form = "column" if "interp_ab_col" in prof else ("gather" if "interp_ab_gather" in prof else "brick")
        return form, (g0, gp, grhs, v, g), gsi

    form, auto, si = run(None)
    forced_form, forced, fsi = run(form)
    say("  chosen: %s" % form)
    assert forced_form == form and si == fsi == 4
    for a, b in zip(auto, forced):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,nb,m", [(64, 1, 1), (128, 2, 2), (128, 3, 0), (64, 3, 1)])
def test_tracers_fp64_columns(B, oracle, monkeypatch, n, nb, m):
    """FP64 boxes of 64^3 and more with n <= 3 run the column walk too (FP64
    planes in the LDS ring): against the brick kernel on the same case and
    against the FP64 oracle, at the FP64 tracer tolerance (1e-12)."""
    from bfps_amd import ic
    say("case: fp64 column walk n=%d nb=%d m=%d" % (n, nb, m))
    npart = 20000
    pos = ic.uniform_positions(npart, seed=40 + nb)
    pos[:20] -= 2 * np.pi
    gp, grhs, gsi, gv, gg = _fp32_tracer_run(B, n, nb, m, pos, "column", monkeypatch, precision="double")
    bp, brhs, bsi, bv, bg = _fp32_tracer_run(B, n, nb, m, pos, "brick", monkeypatch, precision="double")
    say("  column vs brick: pos %.3e rhs %.3e velocity %.3e gradient %.3e"
        % (np.abs(gp - bp).max(), np.abs(grhs - brhs).max(), np.abs(gv - bv).max(), np.abs(gg - bg).max()))
    assert gsi == bsi == 4
    assert np.abs(grhs - brhs).max() <= 1e-13 * max(1.0, np.abs(brhs).max())
    assert np.abs(gp - bp).max() <= 1e-13 * max(1.0, np.abs(bp).max())
    assert np.abs(gv - bv).max() <= 1e-13 * np.abs(bv).max()
    assert np.abs(gg - bg).max() <= 1e-13 * np.abs(bg).max()
    o = oracle.Solver(n, precision="double")
    o.set_physics(0.01, "linear", 1.0, 2.0, 4.0)
    o.set_cfield("vorticity", ic.solenoidal_vorticity(n, seed=21, k0=4.0, dtype=np.complex128))
    o.io_projection()
    op = oracle.Particles(o, npart, nb, m, 4, step_idx=1)
    op.set_state(pos)
    for _ in range(3):
        oracle.nsve_particles_step(o, op, 0.5 / n)
    opos, orhs, osi = op.get_state()
    say("  column vs oracle: pos %.3e rhs %.3e" % (np.abs(gp - opos).max(), np.abs(grhs - orhs).max()))
    assert osi == gsi
    assert np.abs(gp - opos).max() < 1e-12
    assert np.abs(grhs - orhs).max() < 1e-12 * max(1.0, np.abs(orhs).max())


@pytest.mark.parametrize("precision,k,inband", [("double", (1, 2, 0), True), ("double", (2, -1, 3), True),
                                                ("double", (3, 3, 3), False), ("single", (1, 2, 0), True)])
def test_single_wave_rk3_polynomial(B, precision, k, inband):
    """A single Fourier wave has a curl-free u x omega; with linear forcing on
    its shell one IF-SSP-RK3 step multiplies it by
    exp(-nu k^2 dt) (1 + a + a^2/2 + a^3/6), a = A dt."""
    n, nu, dt, A = 32, 0.02, 0.05, 0.5
    say("case: single wave %s %s" % (precision, k))
    kv = np.array(k, float)
    amp = np.array([0.3 + 0.1j, -0.2 + 0.4j, 0.25 - 0.3j])
    amp = amp - kv * (kv @ amp) / (kv @ kv)
    w0 = np.zeros((n, n, n // 2 + 1, 3), dtype=complex)
    w0[k[1] % n, k[2] % n, k[0]] = amp
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.nu, fs.famplitude, fs.forcing_type = nu, A, "linear"
    fs.io_checkpoint(read=True, data=w0)
    fs.step(dt)
    a = A * dt if inband else 0.0
    expect = w0 * np.exp(-nu * (kv @ kv) * dt) * (1 + a + a * a / 2 + a ** 3 / 6)
    e = rel(fs.cvorticity, expect)
    say("  rel", e)
    assert e < (1e-13 if precision == "double" else 1e-6)


def test_kolmogorov_from_rest_simpson(B):
    n, nu, dt, A, fm = 32, 0.1, 0.05, 0.7, 2
    say("case: Kolmogorov from rest")
    fs = B.vorticity_equation("t", n, precision="double")
    fs.nu, fs.famplitude, fs.forcing_type, fs.fmode = nu, A, "Kolmogorov", fm
    fs.io_checkpoint(read=True, data=np.zeros((n, n, n // 2 + 1, 3), dtype=complex))
    lam = nu * fm * fm * dt
    S = (1 + np.exp(-lam)) / 6 + 2 * np.exp(-lam / 2) / 3
    v = 0.0
    expect = np.zeros((n, n, n // 2 + 1, 3), dtype=complex)
    for _ in range(3):
        fs.step(dt)
        v = v * np.exp(-lam) - A / 2 * dt * S
        expect[fm, 0, 0, 2] = v
        expect[n - fm, 0, 0, 2] = v
        assert np.abs(fs.cvorticity - expect).max() < 1e-15


@pytest.mark.parametrize("precision", ["double", "single"])
def test_fixed_energy_injection_vs_oracle(B, oracle, precision):
    """Fixed energy-injection forcing on 1.5 <= |k| <= 3.5 of a random field:
    GPU against the FP64 oracle after 4 steps, and the energy budget of a
    single forced wave (dE/dt = eps - 2 nu k^2 E)."""
    n = 48
    say("case: fixed energy injection %s" % precision)
    w0 = random_vorticity(n, seed=8, dtype=np.complex128)
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = 0.02, 0.3, 1.5, 3.5, "fixed_energy_injection_rate"
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision="double")
    o.set_physics(0.02, "fixed_energy_injection_rate", 0.3, 1.5, 3.5)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    for _ in range(4):
        fs.step(0.01)
        o.step(0.01)
    e = rel(fs.cvorticity, o.get_cfield("vorticity"))
    say("  rel", e)
    assert e < (1e-11 if precision == "double" else 2e-6)


def abc_modes(n, dtype):
    """Spectral ABC vorticity (= velocity) u = (sin z + cos y, sin x + cos z,
    sin y + cos x) in the reference layout, built mode by mode (6 waves)."""
    w = np.zeros((n, n, n // 2 + 1, 3), dtype=dtype)
    # cos(k.x) -> 1/2 at +-k, sin(k.x) -> -i/2 at +k (kx > 0 stored only)
    w[1, 0, 0, 0] += 0.5          # cos y, +ky
    w[n - 1, 0, 0, 0] += 0.5      # cos y, -ky
    w[0, 1, 0, 0] += -0.5j        # sin z, +kz
    w[0, n - 1, 0, 0] += 0.5j     # sin z, -kz
    w[0, 0, 1, 1] += -0.5j        # sin x
    w[0, 1, 0, 1] += 0.5          # cos z
    w[0, n - 1, 0, 1] += 0.5
    w[1, 0, 0, 2] += -0.5j        # sin y
    w[n - 1, 0, 0, 2] += 0.5j
    w[0, 0, 1, 2] += 0.5          # cos x
    return w


@pytest.mark.parametrize("n", [1024])
def test_largest_grid_beltrami_and_tracers(B, n):
    """The 1024^3 FP32 path: exact Beltrami decay of the six modes (and
    nothing else), and tracer velocities against the analytic field."""
    say("case: Beltrami + tracers n=%d single" % n)
    w = abc_modes(n, np.complex64)
    fs = B.vorticity_equation("big", n, precision="single")
    fs.nu, fs.forcing_type = 0.1, "none"
    fs.io_checkpoint(read=True, data=w)
    np_ = 1 << 16
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 2 * np.pi, (np_, 3))
    ps = B.particles_system(fs, np_, 4, 1, 4)
    ps.put(pos)
    fs.velocity_rspace()
    u = ps.sample("velocity")
    ue = np.stack([np.sin(pos[:, 2]) + np.cos(pos[:, 1]), np.sin(pos[:, 0]) + np.cos(pos[:, 2]),
                   np.sin(pos[:, 1]) + np.cos(pos[:, 0])], -1)
    eu = np.abs(u - ue).max()
    say("  tracer velocity max err", eu)
    assert eu < 1e-5
    # dt = 0.5/N as in the workloads: nu k_max^2 dt / 2 stays ~6, so the
    # stage-2 factor exp(+nu k^2 dt/2) of the IF-RK3 (stored in FP32, as in
    # the reference's single mode) cannot overflow
    dt, nsteps = 0.5 / n, 3
    for _ in range(nsteps):
        fs.step(dt)
    got = fs.cvorticity
    idx = np.nonzero(w)
    exact = w[idx].astype(np.complex128) * np.exp(-0.1 * dt * nsteps)
    e_modes = np.abs(got[idx] - exact).max()
    got[idx] = 0
    e_rest = np.abs(got).max()
    say("  mode err %.3e, spurious max %.3e" % (e_modes, e_rest))
    assert len(idx[0]) == 10
    assert e_modes < 1e-6 and e_rest < 1e-6
    del got
    fs.close()


@pytest.mark.parametrize("n,precision,tol", [(48, "double", 1e-14), (64, "single", 1e-6), (32, "double", 1e-14)])
def test_compute_vorticity_vs_oracle(B, oracle, n, precision, tol):
    """compute_vorticity (the curl of cvelocity, cut at kM, symmetrized) of a
    broadband velocity against the oracle; the ABC velocity is its own
    curl."""
    say("case: compute_vorticity n=%d %s" % (n, precision))
    u = random_vorticity(n, seed=31 + n, k0=n / 4.0)
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.compute_vorticity(velocity=u)
    o = oracle.Solver(n, precision=precision)
    o.set_cfield("velocity", u)
    o.compute_vorticity()
    e = rel(fs.cvorticity, o.get_cfield("vorticity"))
    say("  rel vs oracle %.3e" % e)
    assert e < tol
    w = abc_modes(n, fs.cdtype)
    fs.compute_vorticity(velocity=w)
    e = rel(fs.cvorticity, w.astype(np.complex128))
    say("  ABC: curl u = u, rel %.3e" % e)
    assert e < tol
    fs.close()


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-6)])
def test_field_compute_gradient_vs_oracle(B, oracle, precision, tol):
    """compute_gradient(kk, src, dst) on standalone fields (THREE ->
    THREExTHREE) of the velocity, transformed to real space, against the
    oracle's velocity gradient."""
    n = 48
    say("case: field compute_gradient %s" % precision)
    w0 = random_vorticity(n, seed=12)
    fs = B.vorticity_equation("t", n, precision=precision)
    fs.io_checkpoint(read=True, data=w0)
    o = oracle.Solver(n, precision=precision)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    o.gradient_rspace()
    src = B.field(fs, "THREE")
    src.put(fs.compute_velocity(), real_space=False)
    dst = B.field(fs, "THREExTHREE")
    B.compute_gradient(fs.kk, src, dst)
    dst.ift()
    e = rel(dst.get(), o.get_rfield("gradient"))
    say("  rel %.3e" % e)
    assert e < tol
    src.close()
    dst.close()
    fs.close()


def test_impose_zero_modes(B):
    n = 32
    w = random_vorticity(n, seed=13)
    w[0, 0, 0] = [1.0, -2.0, 0.5]
    fs = B.vorticity_equation("t", n, precision="double")
    fs.cvorticity = w
    fs.impose_zero_modes()
    got = fs.cvorticity
    assert np.all(got[0, 0, 0] == 0)
    got[0, 0, 0] = w[0, 0, 0]
    assert np.array_equal(got[..., :n // 2, :], w[..., :n // 2, :])
    fs.close()
