"""ctypes wrapper of the CPU oracle (``liboracle.so``) -- TEST INFRASTRUCTURE ONLY.

Used by ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg as the checker; never by the product library.

Array conventions (the reference's layouts, ``bfps/cpp/field.cpp:85-101``):

* complex fields: ``complex128 [ny][nz][nx//2+1][ncomp]`` (transposed)
* real fields:    ``float64    [nz][ny][nx][ncomp]`` (padding dropped)
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FIELD = {"vorticity": 0, "velocity": 1, "v1": 2, "v2": 3, "rvorticity": 4,
         "rvelocity": 5, "gradient": 6}
FORCING = {"none": 0, "linear": 1, "Kolmogorov": 2, "fixed_energy_injection_rate": 3}

_lib = None


def build(force=False):
    """Compile liboracle.so with g++ (OpenMP)."""
    src = os.path.join(HERE, "bfps_oracle.cpp")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        tmp = LIB + ".tmp"
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
                               src, "-o", tmp])
        os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_new.restype = P
        L.orc_new.argtypes = [I, I, I, D, D, D, I, I]
        L.orc_free.argtypes = [P]
        L.orc_set_threads.argtypes = [I]
        L.orc_get_threads.restype = I
        L.orc_set_physics.argtypes = [P, D, I, D, D, D, I]
        L.orc_kspace_info.argtypes = [P, dp]
        L.orc_kshell.argtypes = [P, dp, ctypes.POINTER(ctypes.c_int64)]
        L.orc_set_cfield.argtypes = [P, I, dp]
        L.orc_get_cfield.argtypes = [P, I, dp]
        L.orc_get_rfield.argtypes = [P, I, dp]
        for name in ("orc_io_projection", "orc_compute_velocity", "orc_compute_vorticity",
                     "orc_velocity_rspace", "orc_gradient_rspace"):
            getattr(L, name).argtypes = [P]
        L.orc_step.argtypes = [P, D]
        L.orc_cospectrum.argtypes = [P, I, I, dp]
        L.orc_pressure.argtypes = [P, dp]
        L.orc_lagrangian_acceleration.argtypes = [P, dp]
        L.orc_rspace_stats.argtypes = [dp, ctypes.c_long, I, I, dp, dp, ctypes.POINTER(ctypes.c_int64)]
        L.orc_fft3d.argtypes = [I, I, I, I, I, I, dp, dp]
        L.orc_symmetrize.argtypes = [I, I, I, I, dp]
        L.orc_interp_table.argtypes = [I, I, dp]
        L.orc_interp_table.restype = I
        L.orc_part_new.restype = P
        L.orc_part_new.argtypes = [P, ctypes.c_long, I, I, I, I]
        L.orc_part_free.argtypes = [P]
        L.orc_part_set.argtypes = [P, dp, dp]
        L.orc_part_get.argtypes = [P, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_part_complete_loop.argtypes = [P, D]
        L.orc_part_sample.argtypes = [P, I, dp]
        L.orc_nsve_particles_step.argtypes = [P, P, D]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def fft3d_r2c(real, fp32=False):
    """real [nz][ny][nx][nc] -> complex [ny][nz][nx//2+1][nc], unnormalised."""
    real = np.ascontiguousarray(real, dtype=np.float64)
    nz, ny, nx, nc = real.shape
    out = np.zeros((ny, nz, nx // 2 + 1, nc), dtype=np.complex128)
    lib().orc_fft3d(nx, ny, nz, nc, int(fp32), 0, _dp(real), _dp(out.view(np.float64)))
    return out


def fft3d_c2r(cplx, nx, fp32=False):
    cplx = np.ascontiguousarray(cplx, dtype=np.complex128)
    ny, nz, nxh, nc = cplx.shape
    out = np.zeros((nz, ny, nx, nc), dtype=np.float64)
    lib().orc_fft3d(nx, ny, nz, nc, int(fp32), 1, _dp(cplx.view(np.float64)), _dp(out))
    return out


def symmetrize(cplx, nx):
    """field::symmetrize of a complex field [ny][nz][nx//2+1][nc] (copy)."""
    out = np.ascontiguousarray(cplx, dtype=np.complex128).copy()
    ny, nz, nxh, nc = out.shape
    lib().orc_symmetrize(nx, ny, nz, nc, _dp(out.view(np.float64)))
    return out


def interp_table(n, m):
    deg = lib().orc_interp_table(n, m, None)
    out = np.zeros(3 * (2 * n + 2) * (deg + 1))
    lib().orc_interp_table(n, m, _dp(out))
    return out.reshape(3, 2 * n + 2, deg + 1)


class Solver:
    """CPU restatement of ``vorticity_equation`` (+ the NSVEparticles step)."""

    def __init__(self, n, precision="double", dealias="SMOOTH", dk=(1.0, 1.0, 1.0), shape=None):
        nx, ny, nz = shape if shape is not None else (n, n, n)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.fp32 = precision in ("single", "float", "fp32")
        self.h = lib().orc_new(nx, ny, nz, dk[0], dk[1], dk[2], int(self.fp32),
                               int(dealias == "SMOOTH"))
        info = np.zeros(4)
        lib().orc_kspace_info(self.h, _dp(info))
        self.kM, self.kM2, self.dk = info[0], info[1], info[2]
        self.nshells = int(info[3])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def set_physics(self, nu, forcing="linear", famplitude=0.5, fk0=2.0, fk1=4.0, fmode=1):
        lib().orc_set_physics(self.h, nu, FORCING[forcing], famplitude, fk0, fk1, fmode)

    def cshape(self, nc=3):
        return (self.ny, self.nz, self.nx // 2 + 1, nc)

    def set_cfield(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128).reshape(self.cshape())
        lib().orc_set_cfield(self.h, FIELD[name], _dp(a.view(np.float64)))

    def get_cfield(self, name):
        nc = 9 if name == "gradient" else 3
        out = np.zeros(self.cshape(nc), dtype=np.complex128)
        lib().orc_get_cfield(self.h, FIELD[name], _dp(out.view(np.float64)))
        return out

    def get_rfield(self, name):
        nc = 9 if name == "gradient" else 3
        out = np.zeros((self.nz, self.ny, self.nx, nc))
        lib().orc_get_rfield(self.h, FIELD[name], _dp(out))
        return out

    def io_projection(self):
        lib().orc_io_projection(self.h)

    def step(self, dt):
        lib().orc_step(self.h, dt)

    def compute_velocity(self):
        lib().orc_compute_velocity(self.h)

    def compute_vorticity(self):
        lib().orc_compute_vorticity(self.h)

    def velocity_rspace(self):
        lib().orc_velocity_rspace(self.h)

    def gradient_rspace(self):
        lib().orc_gradient_rspace(self.h)

    def cospectrum(self, a="velocity", b="velocity"):
        out = np.zeros((self.nshells, 3, 3))
        lib().orc_cospectrum(self.h, FIELD[a], FIELD[b], _dp(out))
        return out

    def pressure(self):
        """compute_pressure of the velocity of cvorticity (spectral, [ny][nz][nx/2+1])."""
        out = np.zeros(self.cshape(1)[:3], dtype=np.complex128)
        lib().orc_pressure(self.h, _dp(out.view(np.float64)))
        return out

    def lagrangian_acceleration(self):
        """compute_Lagrangian_acceleration (spectral, [ny][nz][nx/2+1][3])."""
        out = np.zeros(self.cshape(3), dtype=np.complex128)
        lib().orc_lagrangian_acceleration(self.h, _dp(out.view(np.float64)))
        return out

    def kshell(self):
        ks = np.zeros(self.nshells)
        ns = np.zeros(self.nshells, dtype=np.int64)
        lib().orc_kshell(self.h, _dp(ks), ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return ks, ns


class Particles:
    """CPU restatement of ``particles_system`` on one rank."""

    def __init__(self, solver, np_, n, m, nsteps=4, step_idx=1):
        self.solver = solver
        self.np, self.n, self.m, self.nsteps = int(np_), n, m, nsteps
        self.h = lib().orc_part_new(solver.h, self.np, n, m, nsteps, step_idx)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_part_free(self.h)
            self.h = None

    def set_state(self, pos, rhs=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.np, 3)
        if rhs is None:
            rhs = np.zeros((self.nsteps, self.np, 3))
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        lib().orc_part_set(self.h, _dp(pos), _dp(rhs))

    def get_state(self):
        pos = np.zeros((self.np, 3))
        rhs = np.zeros((self.nsteps, self.np, 3))
        si = ctypes.c_int(0)
        lib().orc_part_get(self.h, _dp(pos), _dp(rhs), ctypes.byref(si))
        return pos, rhs, si.value

    def complete_loop(self, dt):
        lib().orc_part_complete_loop(self.h, dt)

    def sample(self, which="velocity"):
        nc = 9 if which == "gradient" else 3
        out = np.zeros((self.np, nc))
        lib().orc_part_sample(self.h, 1 if which == "velocity" else 2, _dp(out))
        return out


def rspace_stats(data, nbins, max_estimate):
    """field::compute_rspace_stats of real data [..., nc] (nc = 1, 3 or 9):
    moments [10][nvals], histograms [nbins][nvals] (nvals = 4 for nc = 3)."""
    d = np.ascontiguousarray(data, dtype=np.float64)
    nc = d.shape[-1]
    nvals = 4 if nc == 3 else nc
    me = np.ascontiguousarray(max_estimate, dtype=np.float64).reshape(nvals)
    mom = np.zeros((10, nvals))
    hist = np.zeros((nbins, nvals), dtype=np.int64)
    lib().orc_rspace_stats(_dp(d), ctypes.c_long(d.size // nc), nc, nbins, _dp(me), _dp(mom),
                           hist.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return mom, hist


def nsve_particles_step(solver, parts, dt):
    lib().orc_nsve_particles_step(solver.h, parts.h if parts is not None else None, dt)
