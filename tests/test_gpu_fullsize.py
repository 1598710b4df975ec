"""GPU parity at the benchmark configurations' full grid sizes.

* 512^3 (the precision form): one NSVEparticles step of a broadband field
  with 2^16 tracers (n6 m2, AB4) against the CPU oracle, FP64 at rounding
  level and FP32 within 3x the oracle's own FP32-vs-FP64 error; velocity and
  velocity-gradient sampling at the tracers.
* 256^3 FP64 with 2^20 tracers n4 m1 (the FP64 plain form): two steps
  against the FP64 oracle.
* 512^3 and 1024^3 in both precisions, nonlinear term active: a random
  solenoidal field of the modes |k_i| <= 2 stays inside |k_i| <= 16 during
  one IF-RK3 step (each stage's product doubles the band) and no product
  reaches a grid's Nyquist index, so the step on any grid of 128^3 or more
  is the same step; the oracle takes it at 128^3 and the GPU result at
  512^3 / 1024^3 must equal it on those modes and vanish elsewhere.

Tolerances are written in each test.  Each test prints its case before it
starts.
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
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def broadband_vorticity(n, seed, energy=0.5):
    """Random spectral vorticity with |w(k)| ~ 1/(1+|k|) on every stored
    mode (so that every index of every transform carries signal), scaled to
    kinetic energy `energy`; the load projection makes it Hermitian and
    solenoidal and cuts it at kM.  complex64 [ny][nz][nx/2+1][3]."""
    rng = np.random.default_rng(seed)
    nxh = n // 2 + 1
    k1 = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(np.float32)
    kx = np.arange(nxh, dtype=np.float32)
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kx[None, None, :] ** 2
    amp = 1.0 / (1.0 + np.sqrt(k2))
    amp[0, 0, 0] = 0.0
    w = np.empty((n, n, nxh, 3), dtype=np.complex64)
    for c in range(3):
        re = rng.standard_normal((n, n, nxh), dtype=np.float32)
        im = rng.standard_normal((n, n, nxh), dtype=np.float32)
        w[..., c].real = re * amp
        w[..., c].imag = im * amp
    nxm = np.where(kx == 0, 1.0, 2.0).astype(np.float32)
    inv = np.where(k2 > 0, 1.0 / np.maximum(k2, 1.0), 0.0).astype(np.float32)
    e = 0.5 * float(((np.abs(w) ** 2).sum(axis=-1) * inv * nxm[None, None, :]).sum(dtype=np.float64))
    w *= np.float32(np.sqrt(energy / e))
    return w


# ------------------------------------------------------------ 512^3 headline
N512, NU512, DT512, NP512 = 512, 0.0012, 0.5 / 512, 1 << 16


@pytest.fixture(scope="module")
def case512(oracle):
    """Initial field, tracer positions and the FP64 oracle's NSVEparticles
    step of the 512^3 case (shared by the FP64 and FP32 tests)."""
    from bfps_amd import ic
    say("case512: building the broadband 512^3 field and 2^16 tracers")
    w0 = broadband_vorticity(N512, seed=512)
    pos = ic.uniform_positions(NP512, seed=513)
    pos[:64] -= 2 * np.pi  # periodic wrap of negative and large coordinates
    pos[64:128] += 2 * np.pi
    say("case512: FP64 oracle step (n6 m2 tracers)")
    o = oracle.Solver(N512, precision="double")
    o.set_physics(NU512, "linear", 0.5, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    op = oracle.Particles(o, NP512, 6, 2, 4, step_idx=1)
    op.set_state(pos)
    oracle.nsve_particles_step(o, op, DT512)
    v64 = o.get_cfield("vorticity")
    p64, r64, si = op.get_state()
    del op, o
    say("case512: oracle done")
    return {"w0": w0, "pos": pos, "v64": v64, "p64": p64, "r64": r64, "si": si}


def test_512_fp64_step_and_tracers(B, case512):
    """FP64 512^3: vorticity within 1e-11 relative L2, tracer positions and
    velocities within 1e-11 (absolute; velocities are O(1))."""
    say("case: 512^3 FP64 NSVEparticles step vs oracle")
    fs = B.vorticity_equation("t512d", N512, precision="double")
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = NU512, 0.5, 2.0, 4.0, "linear"
    fs.io_checkpoint(read=True, data=case512["w0"])
    ps = B.particles_system(fs, NP512, 6, 2, 4, step_idx=1)
    ps.put(case512["pos"])
    B.NSVEparticles_step(fs, DT512)
    e = rel(fs.cvorticity, case512["v64"])
    gp, grhs, _, gsi = ps.get()
    ep = np.abs(gp - case512["p64"]).max()
    er = np.abs(grhs - case512["r64"]).max()
    say("  vorticity rel %.3e, pos %.3e, rhs %.3e" % (e, ep, er))
    fs.close()
    assert gsi == case512["si"] == 2
    assert e < 1e-11
    assert ep < 1e-11 and er < 1e-11


def test_512_fp32_step_tracers_and_sampling(B, oracle, case512):
    """FP32 512^3 (the benchmark's precision form): vorticity against the
    FP64 oracle within 3x the FP32 oracle's own error (+1e-7); tracer
    velocities against the FP32-field oracle within 1e-5 of their maximum,
    positions within 1e-7; velocity and gradient samples likewise."""
    say("case: 512^3 FP32 NSVEparticles step + samples vs oracle")
    w0, pos = case512["w0"], case512["pos"]
    fs = B.vorticity_equation("t512s", N512, precision="single")
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = NU512, 0.5, 2.0, 4.0, "linear"
    fs.io_checkpoint(read=True, data=w0)
    ps = B.particles_system(fs, NP512, 6, 2, 4, step_idx=1)
    ps.put(pos)
    B.NSVEparticles_step(fs, DT512)
    g32 = fs.cvorticity
    gp, grhs, _, gsi = ps.get()
    fs.velocity_rspace()
    gsv = ps.sample("velocity")
    fs.gradient_rspace()
    gsg = ps.sample("gradient")
    fs.close()
    say("  GPU done; FP32 oracle step")
    o = oracle.Solver(N512, precision="single")
    o.set_physics(NU512, "linear", 0.5, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    op = oracle.Particles(o, NP512, 6, 2, 4, step_idx=1)
    op.set_state(pos)
    oracle.nsve_particles_step(o, op, DT512)
    o32 = o.get_cfield("vorticity")
    opos, orhs, osi = op.get_state()
    o.velocity_rspace()
    osv = op.sample("velocity")
    o.gradient_rspace()
    osg = op.sample("gradient")
    del op, o
    e_gpu = rel(g32, case512["v64"])
    e_orc = rel(o32, case512["v64"])
    say("  vorticity: gpu32 vs orc64 %.3e, orc32 vs orc64 %.3e" % (e_gpu, e_orc))
    er = np.abs(grhs - orhs).max() / np.abs(orhs).max()
    ep = np.abs(gp - opos).max()
    ev = np.abs(gsv - osv).max() / np.abs(osv).max()
    eg = np.abs(gsg - osg).max() / np.abs(osg).max()
    say("  tracers: rhs %.3e pos %.3e; samples: velocity %.3e gradient %.3e" % (er, ep, ev, eg))
    assert gsi == osi == 2
    assert e_gpu < 3 * e_orc + 1e-7
    assert er < 1e-5 and ep < 1e-7
    assert ev < 1e-5 and eg < 1e-5


# ------------------------------------------------------------ 256^3 FP64
def test_256_fp64_tracers_n4m1(B, oracle):
    """The FP64 plain form at its size: 256^3, nu = 0.003, the seeded
    solenoidal field of peak k0 = 4, 2^20 tracers n4 m1 (AB4); two steps
    within 1e-11 (vorticity, relative L2; tracers, absolute)."""
    from bfps_amd import ic
    n, nu, dt, npart = 256, 0.003, 0.5 / 256, 1 << 20
    say("case: 256^3 FP64 + 2^20 tracers n4 m1, 2 steps")
    w0 = ic.solenoidal_vorticity(n, seed=256, k0=4.0, dtype=np.complex128)
    pos = ic.uniform_positions(npart, seed=257)
    fs = B.vorticity_equation("t256", n, precision="double")
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = nu, 0.5, 2.0, 4.0, "linear"
    fs.io_checkpoint(read=True, data=w0)
    ps = B.particles_system(fs, npart, 4, 1, 4, step_idx=1)
    ps.put(pos)
    o = oracle.Solver(n, precision="double")
    o.set_physics(nu, "linear", 0.5, 2.0, 4.0)
    o.set_cfield("vorticity", w0)
    o.io_projection()
    op = oracle.Particles(o, npart, 4, 1, 4, step_idx=1)
    op.set_state(pos)
    for s in range(2):
        B.NSVEparticles_step(fs, dt)
        oracle.nsve_particles_step(o, op, dt)
        say("  step %d done" % s)
    e = rel(fs.cvorticity, o.get_cfield("vorticity"))
    gp, grhs, _, gsi = ps.get()
    opos, orhs, osi = op.get_state()
    ep, er = np.abs(gp - opos).max(), np.abs(grhs - orhs).max()
    say("  vorticity rel %.3e, pos %.3e, rhs %.3e" % (e, ep, er))
    fs.close()
    assert gsi == osi == 3
    assert e < 1e-11
    assert ep < 1e-11 and er < 1e-11


# ------------------------------------------------------------ embedding
EMB_K, EMB_NO, EMB_NU, EMB_DT = 2, 128, 0.002, 0.005


@pytest.fixture(scope="module")
def emb_ref(oracle):
    """The 128^3 FP64 oracle step of the low-mode field, as the |k_i| <= 16
    block of modes."""
    from bfps_amd import ic
    w = ic.solenoidal_vorticity(EMB_NO, seed=77, k0=2.0, kcut=EMB_K, dtype=np.complex128)
    o = oracle.Solver(EMB_NO, precision="double")
    o.set_physics(EMB_NU, "linear", 0.5, 2.0, 4.0)
    o.set_cfield("vorticity", w)
    o.io_projection()
    o.step(EMB_DT)
    ref = o.get_cfield("vorticity")
    kb = 8 * EMB_K
    iy = _low_index(EMB_NO, kb)
    blk = ref[np.ix_(iy, iy, np.arange(kb + 1))].copy()
    # the premise: the 128^3 step leaves nothing outside |k_i| <= 16, and
    # the nonlinear term did act (the result is not the linear decay)
    ref[np.ix_(iy, iy, np.arange(kb + 1))] = 0
    assert np.abs(ref).max() < 1e-14 * np.abs(blk).max()
    w_in = ic.solenoidal_vorticity(EMB_NO, seed=77, k0=2.0, kcut=EMB_K, dtype=np.complex128)
    lin = w_in[np.ix_(iy, iy, np.arange(kb + 1))]
    assert rel(blk, lin) > 1e-3
    return blk


def _low_index(n, kc):
    k = np.arange(-kc, kc + 1)
    return np.where(k >= 0, k, k + n)


@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("precision", ["double", "single"])
def test_embedded_low_mode_step(B, emb_ref, n, precision):
    """One step of a low-mode, non-Beltrami field (u x omega != 0) on the
    full grid equals the 128^3 oracle step on |k_i| <= 16 (FP64 1e-11, FP32
    1e-6 relative L2) and leaves every other mode at rounding level (FP64
    1e-13, FP32 1e-6 of the largest mode)."""
    from bfps_amd import ic
    say("case: embedded low-mode step n=%d %s" % (n, precision))
    kb = 8 * EMB_K
    blk_ref = emb_ref
    fs = B.vorticity_equation("emb", n, precision=precision)
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = EMB_NU, 0.5, 2.0, 4.0, "linear"
    fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=77, k0=2.0, kcut=EMB_K, dtype=fs.cdtype))
    fs.step(EMB_DT)
    got = fs.cvorticity
    fs.close()
    iy = _low_index(n, kb)
    blk = got[np.ix_(iy, iy, np.arange(kb + 1))].astype(np.complex128)
    e_blk = rel(blk, blk_ref)
    # zero the block in place, then the largest remaining mode, row by row
    for a in iy:
        got[a][np.ix_(iy, np.arange(kb + 1))] = 0
    rest = max(float(np.abs(got[y]).max()) for y in range(n))
    del got
    scale = np.abs(blk_ref).max()
    say("  block rel %.3e, largest other mode %.3e of %.3e" % (e_blk, rest, scale))
    if precision == "double":
        assert e_blk < 1e-11 and rest < 1e-13 * scale
    else:
        assert e_blk < 1e-6 and rest < 1e-6 * scale


# ------------------------------------------------------------ fused pass shapes
@pytest.mark.parametrize("n,precision", [(512, "single"), (512, "double"), (1024, "single")])
def test_every_fused_shape_matches_default(B, n, precision):
    """Every compiled tile shape of the step's fused first pass (Biot-Savart
    + inverse z) and last pass (forward y + update) gives the built-in
    shapes' step within rounding (FP64 1e-13, FP32 1e-6 relative L2), on the
    low-mode nonlinear field of the embedding test."""
    from bfps_amd import ic
    fs = B.vorticity_equation("fs", n, precision=precision, plan_rigor="FFTW_ESTIMATE")
    sh = fs.pass_shapes()
    nbs, ny3 = sh["bs"][3], sh["y3"][3]
    say("case: fused shapes n=%d %s: %d first-pass, %d last-pass shapes" % (n, precision, nbs, ny3))
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = EMB_NU, 0.5, 2.0, 4.0, "linear"
    w0 = ic.solenoidal_vorticity(n, seed=77, k0=2.0, kcut=EMB_K, dtype=fs.cdtype)
    tol = 1e-13 if precision == "double" else 1e-6
    iy = _low_index(n, 8 * EMB_K)
    ref = None
    for v in range(max(nbs, ny3)):
        fs.set_pass_shape("bs", min(v, nbs - 1))
        fs.set_pass_shape("y3", min(v, ny3 - 1))
        fs.io_checkpoint(read=True, data=w0)
        fs.step(EMB_DT)
        got = fs.cvorticity
        blk = got[np.ix_(iy, iy, np.arange(8 * EMB_K + 1))].astype(np.complex128)
        rest = float(np.abs(got).sum(dtype=np.float64))
        del got
        if ref is None:
            ref, rest0 = blk, rest
            say("  shapes %s / %s: reference" % (fs.pass_shapes()["bs"][:2], fs.pass_shapes()["y3"][:2]))
            continue
        e = rel(blk, ref)
        say("  shapes %s / %s: block rel %.3e, sum |w| %.6e vs %.6e"
            % (fs.pass_shapes()["bs"][:2], fs.pass_shapes()["y3"][:2], e, rest, rest0))
        assert e < tol and abs(rest - rest0) <= 1e3 * tol * rest0
    fs.close()
