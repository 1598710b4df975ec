"""The launcher side of the NSVEparticles path on one GPU: parameter
defaults and the launch-time rules of bfps/DNS.py, the tracer initial state,
and the driver loop of the generated C++ code (NSVE / NSVEparticles /
direct_numerical_simulation) over this package's device solver.

Files are the reference's (through ``bfps_amd.h5io``): checkpoints
``<simname>_checkpoint_<k>.h5``; the run file ``<simname>.h5`` with the
parameters, ``iteration``/``checkpoint``, the wavenumber shells and the
statistics time series (``write_par``, bfps/_base.py:158-173,
bfps/DNS.py:372-452); the tracer samples ``<simname>_particles.h5``.  When
``<simname>.h5`` is absent the driver keeps the statistics in memory only
(``NSVEparticles.statistics`` holds them in either case).

Differences from files written through libhdf5: the statistics datasets
are contiguous with fixed maximum dimensions rather than chunked and
unlimited (a dataset is grown by appending a longer copy that replaces it
in its group; a statistics row is written in place), and strings are
fixed-length ASCII rather than h5py's variable-length strings.
"""
import os

import numpy as np

from . import __version__
from . import (NSVE_do_stats, NSVEparticles_do_stats, NSVEparticles_read_checkpoint, NSVEparticles_step,
               NSVEparticles_write_checkpoint, h5io, particles_system, vorticity_equation, wisdom_export,
               wisdom_import)


def default_parameters(dns_type="NSVEparticles"):
    """DNS.generate_default_parameters (bfps/DNS.py:139-166), with the
    particle parameters merged for the particle DNS types (DNS.py:655-658)."""
    p = {
        "dealias_type": 1, "dkx": 1.0, "dky": 1.0, "dkz": 1.0,
        "niter_todo": 8, "niter_stat": 1, "niter_out": 8, "checkpoints_per_file": 1,
        "dt": 0.01, "nu": 0.1, "fmode": 1, "famplitude": 0.5, "fk0": 2.0, "fk1": 4.0,
        "forcing_type": "linear", "histogram_bins": 256,
        "max_velocity_estimate": 1.0, "max_vorticity_estimate": 1.0,
    }
    if dns_type in ("NSVEparticles", "NSVEparticles_no_output"):
        p.update({"niter_part": 1, "nparticles": 10, "tracers0_integration_steps": 4,
                  "tracers0_neighbours": 1, "tracers0_smoothness": 1})
    return p


def launch_parameters(n, dns_type="NSVEparticles", precision="single", kMeta=2.0, dtfactor=0.5, **given):
    """The launch-time rules of DNS.prepare_launch (bfps/DNS.py:655-697) for
    an n^3 run: nu = (2 kMeta / n)^(4/3), dt = dtfactor / n, famplitude 0.45
    at n = 288 and 0.47 at n = 576, niter_out = niter_todo when it does not
    divide it, and -- when not given -- as many checkpoints per file as fit
    in 10^9 bytes (one when a single checkpoint is larger).  Values in
    `given` (the command-line parameters) take precedence, as there."""
    p = default_parameters(dns_type)
    p["nx"] = p["ny"] = p["nz"] = int(n)
    p["nu"] = (kMeta * 2 / n) ** (4. / 3)
    p["dt"] = dtfactor / n
    if n == 288:
        p["famplitude"] = 0.45
    elif n == 576:
        p["famplitude"] = 0.47
    cpf = given.pop("checkpoints_per_file", None)
    p.update(given)
    if p["niter_todo"] % p["niter_out"] != 0:
        p["niter_out"] = p["niter_todo"]
    if cpf is None:
        itemsize = 4 if precision == "single" else 8
        size = 3 * (p["nx"] + 2) * p["ny"] * p["nz"] * itemsize
        if "nparticles" in p:
            size += (1 + p["tracers0_integration_steps"]) * 3 * p["nparticles"] * 8
        if size < 1e9:
            cpf = int(1e9 / size)
    if cpf is not None:
        p["checkpoints_per_file"] = int(cpf)
    return p


def kspace_info(p):
    """DNS.get_kspace (bfps/DNS.py:167-192): kM and dk of the dealiasing
    rule, kx (0..nx/2), ky and kz (numpy's fftfreq order times dk), and
    zeroed nshell / kshell of int(kM/dk) + 2 shells (the solver fills
    them in when it starts, kspace::store, bfps/cpp/kspace.cpp:178-230)."""
    div = 2 if p["dealias_type"] == 1 else 3
    kM = max(p["dk" + a] * (p["n" + a] // div - 1) for a in "xyz")
    dk = min(p["dkx"], p["dky"], p["dkz"])
    nshells = int(kM / dk) + 2
    out = {"kM": float(kM), "dk": float(dk), "nshell": np.zeros(nshells, np.int64),
           "kshell": np.zeros(nshells), "kx": np.arange(0, p["nx"] // 2 + 1).astype(np.float64) * p["dkx"]}
    for a in "yz":
        n = p["n" + a]
        k = np.arange(-n // 2 + 1, n // 2 + 1).astype(np.float64) * p["dk" + a]
        out["k" + a] = np.roll(k, n // 2 + 1)
    return out


def write_par(simname, parameters, dns_type="NSVEparticles", iter0=0, work_dir=".", precision="single"):
    """Creates the run file <simname>.h5 as _base.write_par + DNS.write_par
    do (bfps/_base.py:158-173, bfps/DNS.py:372-419): parameters/<name>,
    iteration, bfps_info/{solver_class, exec_name}, kspace/*, one row of
    each statistics dataset (spectra/<f>_<f> [t][nshells][3][3] f64,
    moments/<f> [t][10][4] f64, histograms/<f> [t][nbins][4] i64 for f in
    velocity, vorticity) and checkpoint = 0; for the particle DNS types
    also <simname>_particles.h5 with the tracers0/velocity and
    tracers0/acceleration groups (DNS.launch_jobs, DNS.py:911-916).  The
    schedule assertions of DNS.write_par hold first.  An existing run file
    is an error, as h5py's mode "w-"."""
    p = parameters
    assert p["niter_todo"] % p["niter_stat"] == 0
    assert p["niter_todo"] % p["niter_out"] == 0
    assert p["niter_out"] % p["niter_stat"] == 0
    particles = dns_type in ("NSVEparticles", "NSVEparticles_no_output")
    if particles:
        assert p["niter_todo"] % p["niter_part"] == 0
        assert p["niter_out"] % p["niter_part"] == 0
    os.makedirs(work_dir, exist_ok=True)
    path = os.path.join(work_dir, simname + ".h5")
    if os.path.exists(path):
        raise FileExistsError(path)
    with h5io.File(path, "w") as f:
        for k, v in p.items():
            f["parameters/" + k] = v.encode("ascii") if isinstance(v, str) else v
        f["iteration"] = np.int64(iter0)
        f["bfps_info/solver_class"] = b"DNS"
        # DNS.name (bfps/DNS.py:655): <dns_type>-<precision>-v<version>
        f["bfps_info/exec_name"] = ("%s-%s-v%s" % (dns_type, precision, __version__)).encode("ascii")
        ks = kspace_info(p)
        for k, v in ks.items():
            f["kspace/" + k] = v
        nshells, nbins = ks["nshell"].shape[0], p["histogram_bins"]
        for k in ("velocity", "vorticity"):
            f["statistics/spectra/%s_%s" % (k, k)] = np.zeros((1, nshells, 3, 3))
            f["statistics/moments/" + k] = np.zeros((1, 10, 4))
            f["statistics/histograms/" + k] = np.zeros((1, nbins, 4), np.int64)
        f["checkpoint"] = np.int64(0)
    if particles:
        ppath = os.path.join(work_dir, simname + "_particles.h5")
        if not os.path.exists(ppath):
            with h5io.File(ppath, "w") as f:
                f.create_group("tracers0/velocity")
                f.create_group("tracers0/acceleration")
    return path


def read_iteration(simname, work_dir="."):
    """direct_numerical_simulation::read_iteration
    (bfps/cpp/full_code/direct_numerical_simulation.cpp:17-55): iteration
    and checkpoint from <simname>.h5."""
    with h5io.File(os.path.join(work_dir, simname + ".h5")) as f:
        return int(f["iteration"]), int(f["checkpoint"])


def copy_complex_field(src_path, src_name, dst, dst_name, n, ctype):
    """DNS.copy_complex_field (bfps/DNS.py:810-853): the field
    [nz][ny][nx/2+1][3] of src_path:src_name into dst[dst_name] at n^3, cast
    to ctype; a source of another resolution is cut to, or zero-padded
    into, the leading [min][min][min] block of every axis (the reference's
    rule, which it trusts only for equal factors on all axes).  The
    reference's external link for an equal shape and type becomes a copy."""
    with h5io.File(src_path) as f:
        src = f[src_name]
    shape = (n, n, n // 2 + 1, 3)
    if src.shape == shape:
        dst[dst_name] = src.astype(ctype)
        return
    out = np.zeros(shape, ctype)
    m = tuple(min(a, b) for a, b in zip(shape[:3], src.shape[:3]))
    out[:m[0], :m[1], :m[2]] = src[:m[0], :m[1], :m[2]]
    dst[dst_name] = out


def find_checkpoint(simname, work_dir, iteration):
    """The first <simname>_checkpoint_<k>.h5 holding vorticity/complex/<it>
    (the search of DNS.launch_jobs, bfps/DNS.py:864-874)."""
    k = 0
    while True:
        path = os.path.join(work_dir, "%s_checkpoint_%d.h5" % (simname, k))
        if not os.path.exists(path):
            raise FileNotFoundError("no checkpoint of %s holds iteration %d" % (simname, iteration))
        with h5io.File(path) as f:
            if "vorticity/complex/%d" % iteration in f:
                return path
        k += 1


def launch_jobs(simname, parameters, dns_type="NSVEparticles", work_dir=".", src_simname=None, src_work_dir=".",
                src_iteration=0, particle_rand_seed=None, precision="single"):
    """DNS.launch_jobs (bfps/DNS.py:855-926) up to the job submission: when
    <simname>.h5 does not exist yet, checkpoint 0 gets the initial vorticity
    -- copied from iteration src_iteration of the run src_simname in
    src_work_dir (``--src-simname``, ``--src-wd``, ``--src-iteration``) --
    and, for the particle DNS types, tracers0/state/0 ([np][3]) and
    tracers0/rhs/0 ([nsteps][np][3], zero) with the state drawn by
    generate_tracer_state(particle_rand_seed) (DNS.write_par,
    DNS.py:420-452; DNS.generate_tracer_state, DNS.py:711-733); then the
    run file (write_par) and the tracer-sample file.  Returns the driver,
    not yet initialised.  Without a source run (and no vorticity/complex/0
    in checkpoint 0 yet) the initial vorticity is generate_vector_field with
    the reference's seed 7547, slope 2 and amplitude 0.05."""
    p = parameters
    ckpt0 = os.path.join(work_dir, simname + "_checkpoint_0.h5")
    if not os.path.exists(os.path.join(work_dir, simname + ".h5")):
        os.makedirs(work_dir, exist_ok=True)
        ctype = np.complex64 if precision == "single" else np.complex128
        with h5io.File(ckpt0, "a") as f:
            if "vorticity/complex/0" not in f:
                if src_simname is None:
                    # DNS.launch_jobs without a source run (DNS.py:880-884)
                    f["vorticity/complex/0"] = generate_vector_field(p, spectra_slope=2.0, amplitude=0.05,
                                                                     precision=precision)
                else:
                    src = find_checkpoint(src_simname, src_work_dir, src_iteration)
                    copy_complex_field(src, "vorticity/complex/%d" % src_iteration, f, "vorticity/complex/0",
                                       p["nx"], ctype)
            particles = dns_type in ("NSVEparticles", "NSVEparticles_no_output") and p.get("nparticles", 0) > 0
            if particles:
                np_ = p["nparticles"]
                f["tracers0/rhs/0"] = np.zeros((p["tracers0_integration_steps"], np_, 3))
                f["tracers0/state/0"] = generate_tracer_state(np_, rseed=particle_rand_seed)
        write_par(simname, p, dns_type=dns_type, work_dir=work_dir, precision=precision)
    return NSVEparticles(simname, p, precision=precision, work_dir=work_dir)


def generate_tracer_state(nparticles, rseed=None, batch_size=10 ** 6):
    """DNS.generate_tracer_state (bfps/DNS.py:711-733): uniform positions in
    [0, 2 pi)^3 from numpy's legacy generator (seeded with rseed when
    given), drawn in batches of 10^6 tracers."""
    rand = np.random.RandomState(rseed).random_sample if rseed is not None else np.random.random
    out = np.empty((nparticles, 3))
    done = 0
    while done < nparticles:
        m = min(batch_size, nparticles - done)
        out[done:done + m] = rand((m, 3)) * 2 * np.pi
        done += m
    return out


# ---------------------------------------------------------------- initial fields
def generate_data_3D(n0, n1, n2, dtype=np.complex128, p=1.5, amplitude=0.5, rng=None):
    """tools.generate_data_3D (bfps/tools.py:96-144): a scalar Gaussian
    random field in the FFTW layout [y][z][x/2+1] (n0, n1, n2 the z, y, x
    real-space sizes): amplitude (g1 + i g2) with g1, g2 standard normal
    draws (the real parts first, then the imaginary parts, from numpy's
    legacy global generator, or `rng` -- a RandomState -- when given),
    divided by |k|^p (|k| of the signed wavenumbers, 1 at k = 0); the k = 0
    mode set to 0; the kx = 0 plane made Hermitian in the reference's order
    (ky axis, kz axis, then every (ky, kz) with 1 <= ky < n1/2 taking the
    conjugate of (n1-ky, n0-kz)'s mirror); every mode with |k| > min(n)/3
    set to 0."""
    if n0 % 2 or n1 % 2 or n2 % 2:
        raise ValueError("grid sizes must be even")
    randn = rng.randn if rng is not None else np.random.randn
    shape = (n1, n0, n2 // 2 + 1)
    a = np.zeros(shape, dtype=dtype)
    re = randn(*shape)
    im = randn(*shape)
    a[:] = amplitude * (re + 1j * im)
    # signed wavenumbers: index i <-> i for i <= n/2, i - n above (the
    # reference builds -n/2+1 .. n/2 and rolls it by n/2+1)
    ky = np.where(np.arange(n1) <= n1 // 2, np.arange(n1), np.arange(n1) - n1)
    kz = np.where(np.arange(n0) <= n0 // 2, np.arange(n0), np.arange(n0) - n0)
    kx = np.arange(n2 // 2 + 1)
    k = np.sqrt(ky[:, None, None] ** 2 + kz[None, :, None] ** 2 + kx[None, None, :] ** 2)
    k[0, 0, 0] = 1
    a /= k ** p
    a[0, 0, 0] = 0
    for y in range(1, n1 // 2):
        a[n1 - y, 0, 0] = np.conj(a[y, 0, 0])
    for z in range(1, n0 // 2):
        a[0, n0 - z, 0] = np.conj(a[0, z, 0])
    for y in range(1, n1 // 2):
        a[n1 - y, n0 - np.arange(1, n0), 0] = np.conj(a[y, 1:n0, 0])
    a[k > min(n0, n1, n2) / 3.] = 0
    return a


def generate_data_3D_uniform(n0, n1, n2, dtype=np.complex128, p=1.5, amplitude=1.0, rng=None):
    """tools.generate_data_3D_uniform (bfps/tools.py:61-94): every mode of
    the ky = 0 slab set to `amplitude` (p and rng are ignored)."""
    if n0 % 2 or n1 % 2 or n2 % 2:
        raise ValueError("grid sizes must be even")
    a = np.zeros((n1, n0, n2 // 2 + 1), dtype=dtype)
    a[0] = amplitude
    return a


def padd_with_zeros(a, n0, n1, n2, odtype=None):
    """tools.padd_with_zeros (bfps/tools.py:169-204): the modes of `a`
    ([y][z][x/2+1], possibly with trailing component axes) placed into a
    zero field of real-space size (n0, n1, n2): the four (ky, kz) corner
    blocks of half the source sizes, kx up to m2//2 (m2 = a.shape[2])."""
    odtype = a.dtype if odtype is None else odtype
    if not (a.shape[0] <= n0 and a.shape[1] <= n1 and a.shape[2] <= n2 // 2 + 1):
        raise ValueError("source larger than the target grid")
    b = np.zeros((n0, n1, n2 // 2 + 1) + a.shape[3:], dtype=odtype)
    m0, m1, m2 = a.shape[1], a.shape[0], a.shape[2]
    b[:m1 // 2, :m0 // 2, :m2 // 2 + 1] = a[:m1 // 2, :m0 // 2, :m2 // 2 + 1]
    b[:m1 // 2, n0 - m0 // 2:, :m2 // 2 + 1] = a[:m1 // 2, m0 - m0 // 2:, :m2 // 2 + 1]
    b[n1 - m1 // 2:, :m0 // 2, :m2 // 2 + 1] = a[m1 - m1 // 2:, :m0 // 2, :m2 // 2 + 1]
    b[n1 - m1 // 2:, n0 - m0 // 2:, :m2 // 2 + 1] = a[m1 - m1 // 2:, m0 - m0 // 2:, :m2 // 2 + 1]
    return b


def generate_vector_field(parameters, rseed=7547, spectra_slope=1., amplitude=1., precision="single",
                          scalar_generator=generate_data_3D):
    """DNS.generate_vector_field (bfps/DNS.py:734-809): numpy's legacy
    generator seeded with rseed, three scalar fields of half the grid
    (nz/2, ny/2, nx/2) from `scalar_generator` with slope spectra_slope and
    `amplitude`, cast to the run's complex type, stacked as components and
    padded with zeros to the full grid: [ny][nz][nx/2+1][3].  Not
    divergence free (the load projection of io_checkpoint makes it so)."""
    p = parameters
    ctype = np.complex64 if precision == "single" else np.complex128
    rng = np.random.RandomState(rseed)
    comps = [scalar_generator(p["nz"] // 2, p["ny"] // 2, p["nx"] // 2, p=spectra_slope, amplitude=amplitude,
                              rng=rng).astype(ctype) for _ in range(3)]
    k0 = np.zeros(comps[0].shape + (3,), dtype=comps[0].dtype)
    for c in range(3):
        k0[..., c] = comps[c]
    return padd_with_zeros(k0, p["nz"], p["ny"], p["nx"])


ENSEMBLE_CLASSES = ((32, 32), (32, 48), (24, 64), (8, 128))  # (boxes, N)


def ensemble_manifest(seed=5, classes=ENSEMBLE_CLASSES, nu_range=(0.01, 0.05), k0_choices=(2.0, 3.0, 4.0),
                      np_range=(1e3, 1e5), neighbours=1, smoothness=1, integration_steps=4, dtfactor=0.5):
    """Parameters of an ensemble of independent boxes (no reference
    counterpart: the reference runs one box per job).  One entry per box, in
    the order of `classes` ((count, N) pairs): N, its own initial-field seed,
    nu ~ U[nu_range], peak wavenumber k0 drawn from k0_choices, tracer count
    log-uniform in np_range with its own position seed, dt = dtfactor / N,
    linear band forcing (amplitude 0.5 on 2 <= |k| <= 4, the reference's
    defaults, bfps/DNS.py:139-166) and the tracer scheme (neighbours,
    smoothness, Adams-Bashforth steps; reference defaults 1, 1, 4).  The
    draw is a numpy Generator seeded with `seed`, so a manifest is fixed by
    its arguments."""
    rng = np.random.default_rng(seed)
    boxes = []
    for cnt, n in classes:
        for _ in range(int(cnt)):
            i = len(boxes)
            boxes.append({
                "n": int(n), "seed": int(seed * 1000 + i), "nu": float(rng.uniform(*nu_range)),
                "k0": float(rng.choice(k0_choices)), "dt": dtfactor / n,
                "forcing_type": "linear", "famplitude": 0.5, "fk0": 2.0, "fk1": 4.0,
                "nparticles": int(round(10 ** rng.uniform(np.log10(np_range[0]), np.log10(np_range[1])))),
                "particle_seed": int(seed * 1000 + 500 + i), "neighbours": int(neighbours),
                "smoothness": int(smoothness), "integration_steps": int(integration_steps)})
    return boxes


def write_ensemble_manifest(path, boxes):
    """The manifest as JSON (one object per box)."""
    import json
    with open(path, "w") as f:
        json.dump({"boxes": boxes}, f, indent=1)


def read_ensemble_manifest(path):
    import json
    with open(path) as f:
        return json.load(f)["boxes"]


def build_ensemble(boxes, precision="single", dealias="SMOOTH", with_particles=True):
    """An ``ensemble`` set up from a manifest: each box's physics, its seeded
    solenoidal initial field (``bfps_amd.ic``, peak k0, energy 1/2) through
    the checkpoint-read projections, and its seeded uniform tracers.
    Returns (ensemble, particle systems (None for boxes without tracers),
    dt array)."""
    from . import ensemble, ic
    ens = ensemble([b["n"] for b in boxes], precision=precision, dealias=dealias)
    parts = []
    for i, b in enumerate(boxes):
        fs = ens.boxes[i]
        fs.nu, fs.forcing_type = b["nu"], b["forcing_type"]
        fs.famplitude, fs.fk0, fs.fk1 = b["famplitude"], b["fk0"], b["fk1"]
        fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(b["n"], seed=b["seed"], k0=b["k0"],
                                                                 dtype=fs.cdtype))
        ps = None
        if with_particles and b["nparticles"] > 0:
            ps = particles_system(fs, b["nparticles"], b["neighbours"], b["smoothness"], b["integration_steps"],
                                  step_idx=1)
            ps.put(ic.uniform_positions(b["nparticles"], seed=b["particle_seed"]))
        parts.append(ps)
    return ens, parts, np.array([b["dt"] for b in boxes])


class NSVEparticles:
    """NSVEparticles on one GPU: initialize / step / do_stats /
    write_checkpoint / finalize (bfps/cpp/full_code/NSVE.cpp:7-135,
    NSVEparticles.cpp:7-98) and main_loop
    (direct_numerical_simulation.cpp:88-118, with the stop file of
    code_base.cpp:15-35).  `parameters` is a launch_parameters() dict.
    With `iteration` None the run starts at the iteration and checkpoint
    index stored in <simname>.h5 (as the reference's read_iteration);
    otherwise the checkpoint file holding `iteration` is the one numbered
    `checkpoint`, or searched for when that is None.  When <simname>.h5
    exists, statistics go into it and tracer samples into
    <simname>_particles.h5."""

    def __init__(self, simname, parameters, iteration=None, checkpoint=None, precision="single", work_dir="."):
        self.simname, self.p, self.precision = simname, dict(parameters), precision
        self.work_dir = work_dir
        self.stat_path = self._path(simname + ".h5")
        self.particle_path = self._path(simname + "_particles.h5")
        if iteration is None:
            # the reference starts where <simname>.h5 says
            iteration, checkpoint = read_iteration(simname, work_dir)
        self.iteration0, self.checkpoint0 = int(iteration), checkpoint
        self.statistics = []
        self.fs = self.ps = None

    def _path(self, name):
        return os.path.join(self.work_dir, name)

    def _wisdom_path(self):
        return self._path(self.simname + "_hip_wisdom.txt")

    def initialize(self):
        p = self.p
        name = self._path(self.simname)
        # the transform wisdom of earlier jobs of this run (main_code.hpp:101-104
        # imports <simname>_fftw_wisdom.txt), then the drivers' default rigor
        # (DEFAULT_FFTW_FLAG = FFTW_PATIENT, fftw_interface.hpp:30-35)
        wisdom_import(self._wisdom_path())
        fs = vorticity_equation(name, p["nx"], p["ny"], p["nz"], p["dkx"], p["dky"], p["dkz"],
                                precision=self.precision,
                                dealias="SMOOTH" if p["dealias_type"] == 1 else "TWO_THIRDS",
                                plan_rigor="FFTW_PATIENT")
        fs.nu, fs.fmode, fs.famplitude = p["nu"], p["fmode"], p["famplitude"]
        fs.fk0, fs.fk1, fs.forcing_type = p["fk0"], p["fk1"], p["forcing_type"]
        fs.checkpoints_per_file = p["checkpoints_per_file"]
        fs.iteration = self.iteration0
        if self.checkpoint0 is None:
            k = 0
            while True:
                fs.checkpoint = k
                fname = fs.get_current_fname()
                if not os.path.exists(fname):
                    raise FileNotFoundError("no checkpoint file holds iteration %d" % self.iteration0)
                with h5io.File(fname) as f:
                    if "vorticity/complex/%d" % self.iteration0 in f:
                        break
                k += 1
        else:
            fs.checkpoint = int(self.checkpoint0)
        self.fs = fs
        if "nparticles" in p:
            self.ps = particles_system(fs, p["nparticles"], p["tracers0_neighbours"], p["tracers0_smoothness"],
                                       p["tracers0_integration_steps"], step_idx=fs.iteration + 1)
        NSVEparticles_read_checkpoint(fs, self.ps)
        if os.path.exists(self.stat_path):
            self._grow_and_store_kspace()

    def _grow_and_store_kspace(self):
        """NSVE::initialize on the run file: every statistics dataset grows
        by niter_todo / niter_stat rows (hdf5_tools::grow_file_datasets,
        bfps/cpp/hdf5_tools.cpp:23-90), then kspace::store writes the
        solver's shells (bfps/cpp/kspace.cpp:178-230)."""
        grow = self.p["niter_todo"] // self.p["niter_stat"]
        kk = self.fs.kk
        with h5io.File(self.stat_path, "a") as f:
            for kind in f.keys("statistics"):
                for name in f.keys("statistics/" + kind):
                    key = "statistics/%s/%s" % (kind, name)
                    a = f[key]
                    f[key] = np.concatenate([a, np.zeros((grow,) + a.shape[1:], a.dtype)])
            if f["kspace/kshell"].shape[0] != kk.nshells:
                raise ValueError("run file holds %d shells, the solver %d" % (f["kspace/kshell"].shape[0], kk.nshells))
            f["kspace/kshell"] = kk.kshell
            f["kspace/nshell"] = kk.nshell
            f["kspace/kM"] = np.float64(kk.kM)
            f["kspace/dk"] = np.float64(kk.dk)

    def _write_stats(self, it, st):
        """The statistics rows of field::compute_stats
        (bfps/cpp/field.cpp:836-866, kspace.cpp:561-597) at row
        iteration / niter_stat, and the tracer samples of
        sample_from_particles_system (particles_sampling.hpp:15-50):
        tracers0/<name>/<step_idx> of shape [1][np][3], kept when present."""
        if os.path.exists(self.stat_path) and it % self.p["niter_stat"] == 0:
            row = it // self.p["niter_stat"]
            with h5io.File(self.stat_path, "a") as f:
                for name in ("vorticity", "velocity"):
                    for key, val in (("spectra/%s_%s" % (name, name), st[name]["cospectrum"]),
                                     ("moments/" + name, st[name]["moments"]),
                                     ("histograms/" + name, st[name]["histograms"])):
                        key = "statistics/" + key
                        a = f[key]
                        if row >= a.shape[0]:
                            raise IndexError("%s has %d rows; row %d is past its end" % (key, a.shape[0], row))
                        a[row] = val
                        f[key] = a
        if "tracers0/velocity" in st and os.path.exists(self.particle_path):
            sidx = self.ps.get_step_idx()
            with h5io.File(self.particle_path, "a") as f:
                for name in ("velocity", "acceleration"):
                    key = "tracers0/%s/%d" % (name, sidx)
                    if key not in f:
                        f[key] = st["tracers0/" + name][None]

    @property
    def iteration(self):
        return self.fs.iteration

    def step(self):
        NSVEparticles_step(self.fs, self.p["dt"])

    def do_stats(self):
        """NSVE::do_stats every niter_stat iterations; NSVEparticles adds the
        velocity and Lagrangian acceleration sampled at the tracers every
        niter_part iterations."""
        it, p = self.fs.iteration, self.p
        args = (p["max_velocity_estimate"], p["max_vorticity_estimate"], p["histogram_bins"])
        if self.ps is not None and it % p["niter_part"] == 0:
            st = NSVEparticles_do_stats(self.fs, self.ps, *args)
        elif it % p["niter_stat"] == 0:
            st, tmp = NSVE_do_stats(self.fs, *args)
            tmp.close()
        else:
            return
        self.statistics.append((it, st))
        self._write_stats(it, st)

    def write_checkpoint(self):
        """NSVEparticles::write_checkpoint, then write_iteration
        (direct_numerical_simulation.cpp:57-87): iteration and checkpoint
        into the run file."""
        NSVEparticles_write_checkpoint(self.fs, self.ps)
        if os.path.exists(self.stat_path):
            with h5io.File(self.stat_path, "a") as f:
                f["iteration"] = np.int64(self.fs.iteration)
                f["checkpoint"] = np.int64(self.fs.checkpoint)

    def check_stopping_condition(self):
        return os.path.exists(self._path("stop_" + self.simname))

    def main_loop(self):
        it = self.fs.iteration
        max_iter = it + self.p["niter_todo"] - it % self.p["niter_todo"]
        while self.fs.iteration < max_iter:
            self.do_stats()
            self.step()
            if self.fs.iteration % self.p["niter_out"] == 0:
                self.write_checkpoint()
            if self.check_stopping_condition():
                break
        self.do_stats()
        if self.fs.iteration % self.p["niter_out"] != 0:
            self.write_checkpoint()

    def finalize(self):
        if self.fs is not None:
            self.fs.close()
            self.fs = self.ps = None
            wisdom_export(self._wisdom_path())  # main_code.hpp:145-149


__all__ = ["default_parameters", "launch_parameters", "kspace_info", "write_par", "read_iteration",
           "copy_complex_field", "find_checkpoint", "launch_jobs", "generate_tracer_state", "NSVEparticles",
           "generate_data_3D", "generate_data_3D_uniform", "padd_with_zeros", "generate_vector_field"]
