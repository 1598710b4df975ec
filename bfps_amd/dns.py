"""The launcher side of the NSVEparticles path on one GPU: parameter
defaults and the launch-time rules of bfps/DNS.py, the tracer initial state,
and the driver loop of the generated C++ code (NSVE / NSVEparticles /
direct_numerical_simulation) over this package's device solver.

Statistics are returned (``NSVEparticles.statistics``) instead of being
written to ``<simname>.h5``; checkpoints are the reference's files
(``bfps_amd.h5io``).
"""
import os

import numpy as np

from . import (NSVE_do_stats, NSVEparticles_do_stats, NSVEparticles_read_checkpoint, NSVEparticles_step,
               NSVEparticles_write_checkpoint, h5io, particles_system, vorticity_equation)


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


class NSVEparticles:
    """NSVEparticles on one GPU: initialize / step / do_stats /
    write_checkpoint / finalize (bfps/cpp/full_code/NSVE.cpp:7-135,
    NSVEparticles.cpp:7-98) and main_loop
    (direct_numerical_simulation.cpp:88-118, with the stop file of
    code_base.cpp:15-35).  `parameters` is a launch_parameters() dict.
    The checkpoint file holding `iteration` is found by its index
    (`checkpoint`, as stored in <simname>.h5 by the reference), or searched
    for when that is None."""

    def __init__(self, simname, parameters, iteration=0, checkpoint=None, precision="single", work_dir="."):
        self.simname, self.p, self.precision = simname, dict(parameters), precision
        self.work_dir = work_dir
        self.iteration0, self.checkpoint0 = int(iteration), checkpoint
        self.statistics = []
        self.fs = self.ps = None

    def _path(self, name):
        return os.path.join(self.work_dir, name)

    def initialize(self):
        p = self.p
        name = self._path(self.simname)
        fs = vorticity_equation(name, p["nx"], p["ny"], p["nz"], p["dkx"], p["dky"], p["dkz"],
                                precision=self.precision,
                                dealias="SMOOTH" if p["dealias_type"] == 1 else "TWO_THIRDS")
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
            self.statistics.append((it, NSVEparticles_do_stats(self.fs, self.ps, *args)))
        elif it % p["niter_stat"] == 0:
            st, tmp = NSVE_do_stats(self.fs, *args)
            tmp.close()
            self.statistics.append((it, st))

    def write_checkpoint(self):
        NSVEparticles_write_checkpoint(self.fs, self.ps)

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


__all__ = ["default_parameters", "launch_parameters", "generate_tracer_state", "NSVEparticles"]
