"""Pin the oracle against the reference's golden NSVEparticles checkpoint.

Fixture: bfps/test/B32p1e4_checkpoint_0.h5 (extracted by
tests/golden/extract_fixture.py), asserted by
bfps/test/test_bfps_NSVEparticles.py:41-50 at max-abs 1e-5 for
vorticity/complex/{0,32,64}, tracers0/state/{0,32,64}, tracers0/rhs/{0,32,64}.
Run parameters (test_bfps_NSVEparticles.py:17-34, DNS.py:139-166,660-661):
32^3, single precision fields, nu = (2*2/32)^(4/3), dt = 0.5/32, linear forcing
0.5 on 2<=|k|<=4, 10^4 tracers, neighbours 1, smoothness 1, 4 AB steps,
2 jobs of 32 iterations (restart at 32 re-applies the io_checkpoint
projections and sets step_idx = 33, NSVEparticles.cpp:23).
"""
import numpy as np
import pytest

NU = (2.0 * 2 / 32) ** (4. / 3)
DT = 0.5 / 32


def test_initial_state_is_seeded_uniform(golden):
    np.random.seed(2)
    x = np.random.random((10000, 3)) * 2 * np.pi
    assert np.array_equal(x, golden["state_0"])


@pytest.mark.parametrize("precision,tol", [("single", 1e-7), ("double", 1e-6)])
def test_fixture_two_jobs(oracle, golden, precision, tol):
    s = oracle.Solver(32, precision=precision)
    s.set_physics(NU, "linear", 0.5, 2.0, 4.0)
    s.set_cfield("vorticity", golden["vort_0"])
    s.io_projection()
    p = oracle.Particles(s, 10000, 1, 1, 4, step_idx=1)
    p.set_state(golden["state_0"], golden["rhs_0"])
    for _ in range(32):
        oracle.nsve_particles_step(s, p, DT)
    v = s.get_cfield("vorticity")
    pos, rhs, si = p.get_state()
    assert si == 33
    assert np.abs(v - golden["vort_32"]).max() < tol
    assert np.abs(pos - golden["state_32"]).max() < tol
    assert np.abs(rhs - golden["rhs_32"]).max() < 10 * tol
    # second job: restart from the written checkpoint
    vs = v.astype(np.complex64) if precision == "single" else v
    s2 = oracle.Solver(32, precision=precision)
    s2.set_physics(NU, "linear", 0.5, 2.0, 4.0)
    s2.set_cfield("vorticity", vs)
    s2.io_projection()
    p2 = oracle.Particles(s2, 10000, 1, 1, 4, step_idx=33)
    p2.set_state(pos, rhs)
    for _ in range(32):
        oracle.nsve_particles_step(s2, p2, DT)
    v = s2.get_cfield("vorticity")
    pos, rhs, si = p2.get_state()
    assert si == 65
    assert np.abs(v - golden["vort_64"]).max() < tol
    assert np.abs(pos - golden["state_64"]).max() < tol
    assert np.abs(rhs - golden["rhs_64"]).max() < 10 * tol


def test_fixture_tolerance_is_meaningful(oracle, golden):
    """A dropped forcing term must fail the reference's 1e-5 criterion."""
    s = oracle.Solver(32, precision="single")
    s.set_physics(NU, "none")
    s.set_cfield("vorticity", golden["vort_0"])
    s.io_projection()
    for _ in range(32):
        s.step(DT)
    assert np.abs(s.get_cfield("vorticity") - golden["vort_32"]).max() > 1e-5
