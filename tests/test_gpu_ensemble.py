"""GPU tests of the batched ensemble: boxes of uneven sizes and physics
advanced by one graph launch per step must match the same boxes advanced
one by one (same kernels: fields bit for bit; tracers within rounding of the
two interpolation kernels), on every replay of the captured graph."""
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


PHYS = [  # nu, forcing, famplitude, fk0, fk1, fmode
    (0.02, "linear", 0.5, 2.0, 4.0, 1),
    (0.03, "none", 0.5, 2.0, 4.0, 1),
    (0.015, "Kolmogorov", 0.4, 2.0, 4.0, 2),
    (0.025, "fixed_energy_injection_rate", 0.2, 1.5, 3.5, 1),
    (0.04, "linear", 0.3, 1.0, 3.0, 1),
]


def setup_box(fs, i, n, ic, seed):
    nu, forcing, famp, fk0, fk1, fmode = PHYS[i % len(PHYS)]
    fs.nu, fs.forcing_type, fs.famplitude, fs.fk0, fs.fk1, fs.fmode = nu, forcing, famp, fk0, fk1, fmode
    fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=seed, k0=2.0 + i % 3,
                                                             dtype=fs.cdtype))


@pytest.mark.parametrize("precision", ["single", "double"])
def test_ensemble_matches_separate_boxes(B, precision):
    from bfps_amd import ic
    sizes = [32, 48, 32, 64, 16, 48]
    say("case: ensemble %s sizes %s" % (precision, sizes))
    ens = B.ensemble(sizes, precision=precision)
    ref = [B.vorticity_equation("r%d" % i, n, precision=precision) for i, n in enumerate(sizes)]
    npart = {0: 500, 3: 2000, 5: 1000}
    ps_e, ps_r = {}, {}
    for i, n in enumerate(sizes):
        setup_box(ens.boxes[i], i, n, ic, seed=10 + i)
        setup_box(ref[i], i, n, ic, seed=10 + i)
        if i in npart:
            pos = ic.uniform_positions(npart[i], seed=100 + i)
            ps_e[i] = B.particles_system(ens.boxes[i], npart[i], 2, 1, 4, step_idx=1)
            ps_e[i].put(pos)
            ps_r[i] = B.particles_system(ref[i], npart[i], 2, 1, 4, step_idx=1)
            ps_r[i].put(pos)
    dt = np.array([0.5 / n for n in sizes])
    for step in range(4):
        ens.step(dt, with_particles=True)
        for i in range(len(sizes)):
            B.NSVEparticles_step(ref[i], dt[i])
        ens.synchronize()
        for i in range(len(sizes)):
            a, b = ens.boxes[i].cvorticity, ref[i].cvorticity
            assert np.array_equal(a, b), (step, i, np.abs(a - b).max())
        for i in npart:
            pa, pb = ps_e[i].get()[0], ps_r[i].get()[0]
            # brick kernel (FP32 inner sums) against the global-gather kernel (FP64)
            tol = 1e-12 if precision == "double" else 1e-7
            assert np.abs(pa - pb).max() < tol, (step, i, np.abs(pa - pb).max())
        say("  step %d: fields identical, tracers agree" % step)
    assert all(b.iteration == 4 for b in ens.boxes)
    # a change of physics and time step is picked up without recapture errors
    ens.boxes[1].nu = 0.05
    ref[1].nu = 0.05
    dt2 = dt * 0.5
    ens.step(dt2, with_particles=False)
    for i in range(len(sizes)):
        ref[i].step(dt2[i])
    for i in range(len(sizes)):
        assert np.array_equal(ens.boxes[i].cvorticity, ref[i].cvorticity), i
    # a box of the 48^3 class (no fixed-injection box so far) switches to
    # that forcing: the graph must gain the band-amplitude launches
    for f in (ens.boxes[1], ref[1]):
        f.forcing_type, f.famplitude, f.fk0, f.fk1 = "fixed_energy_injection_rate", 0.2, 1.5, 3.5
    for step in range(2):
        ens.step(dt2, with_particles=False)
        for i in range(len(sizes)):
            ref[i].step(dt2[i])
        for i in range(len(sizes)):
            assert np.array_equal(ens.boxes[i].cvorticity, ref[i].cvorticity), (step, i)
    say("  forcing switched to fixed injection: fields identical")
    ens.close()


def test_ensemble_errors(B):
    with pytest.raises(B.capi.BfpshError):
        B.ensemble([32, 40])
    with pytest.raises(B.capi.BfpshError):
        B.ensemble([])


def test_ensemble_full_manifest(B, oracle):
    """The 96-box ensemble at its full size (32 x 32^3, 32 x 48^3, 24 x 64^3,
    8 x 128^3; per-box nu, k0 and log-uniform tracer counts, n1 m1 AB4):
    three batched graph steps against the same manifest advanced box by box
    (same kernels: fields and tracers bit for bit), and the first 128^3 box
    against the FP32 oracle (vorticity within 3e-6 relative L2, tracer
    positions within 1e-7, velocities within 1e-5 of their maximum)."""
    from bfps_amd import dns, ic
    boxes = dns.ensemble_manifest(seed=5)
    assert len(boxes) == 96 and sum(b["n"] ** 3 for b in boxes) == 27656192
    say("case: 96-box ensemble, %d tracers" % sum(b["nparticles"] for b in boxes))
    ens, pe, dt = dns.build_ensemble(boxes)
    sep, psep, _ = dns.build_ensemble(boxes)
    nsteps = 3
    for s in range(nsteps):
        ens.step(dt)
        sep.step_separate(dt)
    ens.synchronize()
    sep.synchronize()
    for i in range(len(boxes)):
        a, b = ens.boxes[i].cvorticity, sep.boxes[i].cvorticity
        assert np.array_equal(a, b), (i, np.abs(a - b).max())
        if pe[i] is not None:
            assert np.array_equal(pe[i].get()[0], psep[i].get()[0]), i
    say("  batched == box by box for all 96 boxes")
    i = next(k for k, b in enumerate(boxes) if b["n"] == 128)
    b = boxes[i]
    o = oracle.Solver(128, precision="single")
    o.set_physics(b["nu"], "linear", b["famplitude"], b["fk0"], b["fk1"])
    o.set_cfield("vorticity", ic.solenoidal_vorticity(128, seed=b["seed"], k0=b["k0"]))
    o.io_projection()
    op = oracle.Particles(o, b["nparticles"], b["neighbours"], b["smoothness"], b["integration_steps"], step_idx=1)
    op.set_state(ic.uniform_positions(b["nparticles"], seed=b["particle_seed"]))
    for s in range(nsteps):
        oracle.nsve_particles_step(o, op, b["dt"])
    ref = o.get_cfield("vorticity")
    e = np.linalg.norm(ens.boxes[i].cvorticity - ref) / np.linalg.norm(ref)
    gp, grhs, _, gsi = pe[i].get()
    opos, orhs, osi = op.get_state()
    ep, er = np.abs(gp - opos).max(), np.abs(grhs - orhs).max() / np.abs(orhs).max()
    say("  box %d (128^3, %d tracers) vs oracle: vorticity %.3e pos %.3e rhs %.3e"
        % (i, b["nparticles"], e, ep, er))
    ens.close()
    sep.close()
    assert gsi == osi == nsteps + 1
    assert e < 3e-6 and ep < 1e-7 and er < 1e-5
