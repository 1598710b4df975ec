"""Launcher semantics (bfps/DNS.py) restated in bfps_amd/dns.py, pinned by
the reference's golden run: its tracer initial state is the seeded legacy
numpy draw (--particle-rand-seed 2), and its physics (nu, dt, forcing,
interpolation and integration settings) are the launch rules applied to
n = 32 -- the GPU test that runs the golden case through the driver checks
the trajectory they produce at the reference's 1e-5."""
import numpy as np

import bfps_amd

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


def test_write_par_run_file(tmp_path):
    # the run file of DNS.write_par (bfps/_base.py:158-173, bfps/DNS.py:372-419)
    import os
    import sys
    import pytest
    from bfps_amd import h5io
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import h5mini
    p = dns.launch_parameters(32, nparticles=10000, niter_todo=32, niter_out=32, niter_stat=16, niter_part=16,
                              histogram_bins=64)
    path = dns.write_par("run", p, work_dir=str(tmp_path))
    assert dns.read_iteration("run", str(tmp_path)) == (0, 0)
    with h5io.File(path) as f:
        assert f["parameters/nu"] == NU and f["parameters/nu"].dtype == np.float64
        assert f["parameters/nx"] == 32 and f["parameters/nx"].dtype == np.int64
        assert f["parameters/forcing_type"] == b"linear"
        assert f["bfps_info/exec_name"] == ("NSVEparticles-single-v%s" % bfps_amd.__version__).encode()
        # wavenumbers: k_i = i up to the Nyquist index (positive), i - n above
        n = 32
        signed = np.array([i if i <= n // 2 else i - n for i in range(n)], dtype=np.float64)
        assert np.array_equal(f["kspace/kx"], np.arange(n // 2 + 1.0))
        assert np.array_equal(f["kspace/ky"], signed) and np.array_equal(f["kspace/kz"], signed)
        assert f["kspace/kM"] == 15.0 and f["kspace/dk"] == 1.0
        assert f["kspace/nshell"].shape == (17,) and f["kspace/nshell"].dtype == np.int64
        for k in ("velocity", "vorticity"):
            assert f["statistics/spectra/%s_%s" % (k, k)].shape == (1, 17, 3, 3)
            assert f["statistics/moments/" + k].shape == (1, 10, 4)
            assert f["statistics/histograms/" + k].shape == (1, 64, 4)
            assert f["statistics/histograms/" + k].dtype == np.int64
        assert f.keys("statistics") == ["histograms", "moments", "spectra"]
    # the independent test reader agrees
    assert h5mini.read(path, "parameters/forcing_type") == b"linear"
    assert h5mini.read(path, "statistics/moments/velocity").shape == (1, 10, 4)
    with h5io.File(os.path.join(str(tmp_path), "run_particles.h5")) as f:
        assert f.keys("tracers0") == ["acceleration", "velocity"]
    with pytest.raises(FileExistsError):
        dns.write_par("run", p, work_dir=str(tmp_path))
    # TWO_THIRDS shells: kM = dk (n // 3 - 1)
    q = dns.launch_parameters(48, dns_type="NSVE", dealias_type=0, niter_todo=8, niter_out=8)
    ks = dns.kspace_info(q)
    assert ks["kM"] == 15.0 and ks["nshell"].shape == (17,)
    # the schedule rules of DNS.write_par
    bad = dict(p, niter_out=8, niter_stat=3)
    with pytest.raises(AssertionError):
        dns.write_par("bad", bad, work_dir=str(tmp_path))


def test_launch_jobs_from_source_run(tmp_path, golden):
    # DNS.launch_jobs with --src-simname B32p1e4 --src-iteration 0
    # --particle-rand-seed 2 (the reference's test_bfps_NSVEparticles.py):
    # checkpoint 0 holds the source field, the seeded tracer state and a
    # zero history; the run file says iteration 0, checkpoint 0
    import os
    from bfps_amd import h5io
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    p = dns.launch_parameters(32, nparticles=10000, niter_todo=32, niter_out=32, niter_stat=1,
                              checkpoints_per_file=3)
    run = dns.launch_jobs("dns_nsveparticles", p, work_dir=str(tmp_path), src_simname="B32p1e4",
                          src_work_dir=gold, src_iteration=0, particle_rand_seed=2)
    assert (run.iteration0, run.checkpoint0) == (0, 0)
    with h5io.File(os.path.join(str(tmp_path), "dns_nsveparticles_checkpoint_0.h5")) as f:
        w = f["vorticity/complex/0"]
        assert w.dtype == np.complex64 and np.array_equal(w, golden["vort_0"])
        assert np.array_equal(f["tracers0/state/0"], golden["state_0"])
        assert f["tracers0/rhs/0"].shape == (4, 10000, 3) and not f["tracers0/rhs/0"].any()
    # a second launch of the same run changes nothing (the run file exists)
    dns.launch_jobs("dns_nsveparticles", p, work_dir=str(tmp_path))
    # another resolution: the source's leading block, zero elsewhere
    q = dns.launch_parameters(48, nparticles=10, niter_todo=8, niter_out=8)
    dns.launch_jobs("up", q, work_dir=str(tmp_path), src_simname="B32p1e4", src_work_dir=gold,
                    particle_rand_seed=1)
    with h5io.File(os.path.join(str(tmp_path), "up_checkpoint_0.h5")) as f:
        w = f["vorticity/complex/0"]
    assert w.shape == (48, 48, 25, 3)
    assert np.array_equal(w[:32, :32, :17], golden["vort_0"]) and not w[32:].any() and not w[:, 32:].any()
    # without a source: DNS.generate_vector_field(spectra_slope=2, amplitude=0.05),
    # seed 7547 (DNS.py:880-884)
    dns.launch_jobs("gen", q, work_dir=str(tmp_path / "x"), particle_rand_seed=1)
    with h5io.File(os.path.join(str(tmp_path / "x"), "gen_checkpoint_0.h5")) as f:
        w = f["vorticity/complex/0"]
    want = dns.generate_vector_field(q, spectra_slope=2.0, amplitude=0.05)
    assert w.dtype == np.complex64 and np.array_equal(w, want) and np.abs(w).max() > 0


def _signed(n):
    return np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)


def test_generate_data_3D_structure():
    """tools.generate_data_3D (bfps/tools.py:96-144): the draws of numpy's
    legacy generator (real parts, then imaginary parts) over |k|^p on the
    modes the symmetrisation and the |k| > n/3 cut leave alone; the k = 0
    mode zero; the kx = 0 plane Hermitian; nothing beyond n/3.  (Exact
    values are not pinned by a reference vector: bfps.tools imports h5py,
    which is absent here.)"""
    n0, n1, n2, p, amp = 12, 16, 20, 1.5, 0.5
    a = dns.generate_data_3D(n0, n1, n2, p=p, amplitude=amp, rng=np.random.RandomState(3))
    assert a.shape == (n1, n0, n2 // 2 + 1) and a.dtype == np.complex128
    rs = np.random.RandomState(3)
    g = amp * (rs.randn(n1, n0, n2 // 2 + 1) + 1j * rs.randn(n1, n0, n2 // 2 + 1))
    ky, kz, kx = _signed(n1), _signed(n0), np.arange(n2 // 2 + 1)
    k = np.sqrt(ky[:, None, None] ** 2 + kz[None, :, None] ** 2 + kx[None, None, :] ** 2)
    cut = k > min(n0, n1, n2) / 3.
    assert not a[cut].any() and a[0, 0, 0] == 0
    # modes outside the kx = 0 plane's mirrored half keep their draw
    keep = ~cut
    keep[:, :, 0] = False
    kk = k.copy()
    kk[0, 0, 0] = 1
    assert np.allclose(a[keep], (g / kk ** p)[keep], rtol=1e-15, atol=0)
    # Hermitian kx = 0 plane (inside the cut)
    for y in range(n1):
        for z in range(n0):
            if not cut[y, z, 0]:
                assert a[(n1 - y) % n1, (n0 - z) % n0, 0] == np.conj(a[y, z, 0]) or (y in (0, n1 // 2) and z in (0, n0 // 2))
    # the global generator is used when no RandomState is given
    np.random.seed(3)
    assert np.array_equal(dns.generate_data_3D(n0, n1, n2, p=p, amplitude=amp), a)


def test_padd_with_zeros_and_vector_field():
    # (the reference pads cubic grids: it mixes the y and z sizes of
    # non-cubic ones, bfps/tools.py:196-203)
    a = np.arange(6 * 6 * 4 * 2).reshape(6, 6, 4, 2).astype(np.complex64) + 1
    b = dns.padd_with_zeros(a, 10, 10, 12)
    assert b.shape == (10, 10, 7, 2) and b.dtype == np.complex64
    # the four (ky, kz) corner blocks of half the source sizes, kx < 3
    assert np.array_equal(b[:3, :3, :3], a[:3, :3, :3]) and np.array_equal(b[-3:, -3:, :3], a[-3:, -3:, :3])
    assert np.array_equal(b[:3, -3:, :3], a[:3, -3:, :3]) and np.array_equal(b[-3:, :3, :3], a[-3:, :3, :3])
    assert b.astype(bool).sum() == 4 * 3 * 3 * 3 * 2
    # vector field: three seeded components of the half grid, padded
    q = dns.launch_parameters(32)
    v = dns.generate_vector_field(q, rseed=11, spectra_slope=2.0, amplitude=0.05)
    assert v.shape == (32, 32, 17, 3) and v.dtype == np.complex64
    assert np.array_equal(v, dns.generate_vector_field(q, rseed=11, spectra_slope=2.0, amplitude=0.05))
    rs = np.random.RandomState(11)
    comps = [dns.generate_data_3D(16, 16, 16, p=2.0, amplitude=0.05, rng=rs).astype(np.complex64) for _ in range(3)]
    for c in range(3):
        assert np.array_equal(v[..., c], dns.padd_with_zeros(comps[c], 32, 32, 32))
    ky = _signed(32)
    assert not v[np.abs(ky) >= 8].any() and not v[:, np.abs(ky) >= 8].any() and not v[:, :, 9:].any()
    assert not np.array_equal(v[..., 0], v[..., 1])
    u = dns.generate_vector_field(q, rseed=11, precision="double", scalar_generator=dns.generate_data_3D_uniform)
    assert u.dtype == np.complex128 and np.all(u[0, :8, :5] == 1.0) and u.astype(bool).sum() == 3 * 16 * 5
