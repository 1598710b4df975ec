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
    ens.close()


def test_ensemble_errors(B):
    with pytest.raises(B.capi.BfpshError):
        B.ensemble([32, 40])
    with pytest.raises(B.capi.BfpshError):
        B.ensemble([])
