"""Launcher semantics (bfps/DNS.py) restated in bfps_amd/dns.py, pinned by
the reference's golden run: its tracer initial state is the seeded legacy
numpy draw (--particle-rand-seed 2), and its physics (nu, dt, forcing,
interpolation and integration settings) are the launch rules applied to
n = 32 -- the GPU test that runs the golden case through the driver checks
the trajectory they produce at the reference's 1e-5."""
import numpy as np

from bfps_amd import dns

NU = (2.0 * 2 / 32) ** (4. / 3)
DT = 0.5 / 32


def test_tracer_state_is_the_golden_initial_state(golden):
    assert np.array_equal(dns.generate_tracer_state(10000, rseed=2), golden["state_0"])
    # batches of 10^6 draw the same stream as one draw
    a = dns.generate_tracer_state(2500, rseed=9, batch_size=1000)
    assert np.array_equal(a, dns.generate_tracer_state(2500, rseed=9))


def test_launch_parameters_of_the_golden_run():
    p = dns.launch_parameters(32, nparticles=10000, niter_todo=32, niter_out=32, niter_stat=1,
                              checkpoints_per_file=3)
    assert p["nu"] == NU and p["dt"] == DT
    assert (p["famplitude"], p["fk0"], p["fk1"], p["forcing_type"]) == (0.5, 2.0, 4.0, "linear")
    assert (p["tracers0_neighbours"], p["tracers0_smoothness"], p["tracers0_integration_steps"]) == (1, 1, 4)
    assert p["checkpoints_per_file"] == 3 and p["dealias_type"] == 1


def test_launch_rules():
    # checkpoints per file: 10^9 bytes over 3 (nx+2) ny nz e + (1 + nsteps) 3 np 8
    p = dns.launch_parameters(32, nparticles=10000)
    assert p["checkpoints_per_file"] == int(1e9 / (3 * 34 * 32 * 32 * 4 + 5 * 3 * 10000 * 8))
    assert dns.launch_parameters(1024, precision="double")["checkpoints_per_file"] == 1
    assert dns.launch_parameters(288)["famplitude"] == 0.45
    assert dns.launch_parameters(576)["famplitude"] == 0.47
    assert dns.launch_parameters(64, niter_todo=10, niter_out=8)["niter_out"] == 10
    assert dns.launch_parameters(64, nu=0.02)["nu"] == 0.02          # a given value wins
    assert "nparticles" not in dns.launch_parameters(64, dns_type="NSVE")
