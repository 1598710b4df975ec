"""bfps_amd -- MI355X (gfx950) backend for bfps's vorticity-equation step and
tracer loop.

Python mirror of the reference's class surface over the C ABI
(include/bfpsh.h):

* ``vorticity_equation``  <- ``vorticity_equation<rnumber, FFTW>``
  (bfps/cpp/vorticity_equation.hpp:46-134)
* ``kspace``              <- ``kspace<FFTW, SMOOTH>`` metadata
  (bfps/cpp/kspace.hpp:43-195)
* ``particles_system``    <- ``abstract_particles_system`` built by
  ``particles_system_builder`` (bfps/cpp/particles/abstract_particles_system.hpp:12-55)
* ``field``               <- ``field<rnumber, FFTW, ONE|THREE|THREExTHREE>``
  with its in-place ``dft``/``ift``/``symmetrize`` (bfps/cpp/field.hpp:47-200)
* ``NSVEparticles_step``  <- ``NSVEparticles::step`` (bfps/cpp/full_code/NSVEparticles.cpp:33-41)
* ``ensemble``            -- many independent boxes advanced together (one
  graph launch of batched field kernels per step); no reference counterpart

Host arrays use the reference's layouts (complex ``[ny][nz][nx//2+1][3]``,
real ``[nz][ny][nx][ncomp]``).  Everything runs on the GPU through
libbfps_hip.so; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import capi
from . import h5io

__version__ = "0.2.0"

__all__ = ["vorticity_equation", "kspace", "field", "compute_gradient", "particles_system", "NSVEparticles_step",
           "ensemble", "capi"]


class kspace:
    """Fourier-space geometry of a context (bfps/cpp/kspace.cpp:34-167)."""

    def __init__(self, ctx, nx, ny, nz, dkx, dky, dkz):
        lib = capi.load()
        info = np.zeros(4)
        capi.check(ctx, lib.bfpsh_kspace_info(ctx, info.ctypes.data_as(capi._DP), None, None), "kspace_info")
        self.kM, self.kM2, self.dk = float(info[0]), float(info[1]), float(info[2])
        self.nshells = int(info[3])
        self.kshell = np.zeros(self.nshells)
        self.nshell = np.zeros(self.nshells, dtype=np.int64)
        capi.check(ctx, lib.bfpsh_kspace_info(ctx, info.ctypes.data_as(capi._DP),
                                              self.kshell.ctypes.data_as(capi._DP),
                                              self.nshell.ctypes.data_as(capi._LP)), "kspace_info")
        self.dkx, self.dky, self.dkz = dkx, dky, dkz
        self.dk2 = self.dk ** 2
        self.kx = np.arange(nx // 2 + 1) * dkx
        self.ky = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny) * dky
        self.kz = np.where(np.arange(nz) <= nz // 2, np.arange(nz), np.arange(nz) - nz) * dkz


class vorticity_equation:
    """GPU-resident ``vorticity_equation<rnumber, HIP>``.

    precision:  'single' (rnumber = float) or 'double'.
    dealias:    'SMOOTH' (the reference solver's fixed choice) or 'TWO_THIRDS'.
    plan_rigor: 'FFTW_ESTIMATE', 'FFTW_MEASURE' (the reference constructor's
                default) or 'FFTW_PATIENT': how the strided FFT passes pick
                their tile shapes (``set_plan_rigor``).
    """

    def __init__(self, name, nx, ny=None, nz=None, dkx=1.0, dky=1.0, dkz=1.0,
                 precision="single", dealias="SMOOTH", device=0, plan_rigor="FFTW_MEASURE", _handle=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.name = name
        self.nx, self.ny, self.nz = nx, ny, nz
        self.precision = precision
        self.rdtype = np.float32 if precision == "single" else np.float64
        self.cdtype = np.complex64 if precision == "single" else np.complex128
        lib = capi.load()
        # _handle: a box context owned by an ensemble (not destroyed here)
        self._owned = _handle is None
        if _handle is not None:
            h = ctypes.c_void_p(_handle)
        else:
            h = ctypes.c_void_p()
            rc = lib.bfpsh_create(ctypes.byref(h), nx, ny, nz, dkx, dky, dkz,
                                  capi.FP32 if precision == "single" else capi.FP64,
                                  capi.SMOOTH if dealias == "SMOOTH" else capi.TWO_THIRDS, device)
            if rc != 0:
                msg = lib.bfpsh_last_error(h)
                lib.bfpsh_destroy(h)
                self._h = None
                raise capi.BfpshError("bfpsh_create failed (%d): %s" % (rc, msg.decode() if msg else ""))
        self._h = h
        self.kk = kspace(h, nx, ny, nz, dkx, dky, dkz)
        if self._owned:
            self.set_plan_rigor(plan_rigor)
        # physical parameters, initialised as the reference constructor does
        # (bfps/cpp/vorticity_equation.cpp:149-155)
        self.nu = 0.1
        self.fmode = 1
        self.famplitude = 1.0
        self.fk0 = 2.0
        self.fk1 = 4.0
        self.forcing_type = "linear"
        self.particles = None
        # checkpoint files <name>_checkpoint_<checkpoint>.h5, holding up to
        # checkpoints_per_file fields (NSVE.cpp:46-56 sets both)
        self.checkpoint = 0
        self.checkpoints_per_file = 1

    # ---------------------------------------------------------------- basics
    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                capi.load().bfpsh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_plan_rigor(self, rigor):
        """Choose the tile shape of each strided FFT pass (the reference's
        plan rigor, bfps/cpp/field.cpp:105-118): 'FFTW_ESTIMATE' takes the
        wisdom or the built-in shapes, 'FFTW_MEASURE' / 'FFTW_PATIENT' time
        the compiled shapes unless the wisdom holds a choice of that rigor."""
        if rigor not in capi.PLAN_RIGOR:
            raise capi.BfpshError("unknown plan rigor %r" % (rigor,))
        capi.check(self._h, capi.load().bfpsh_set_plan_rigor(self._h, capi.PLAN_RIGOR[rigor]), "set_plan_rigor")

    def pass_shapes(self):
        """{kind: (values per thread, columns, shape index, shapes compiled,
        persistent)} of the strided passes (kinds z_inv, y_inv, z_fwd,
        y_fwd, and the step's fused bs and y3)."""
        out = {}
        for name, k in capi.PASS_KINDS.items():
            e, b, pp, idx, n = [ctypes.c_int() for _ in range(5)]
            capi.check(self._h, capi.load().bfpsh_get_pass_shape(self._h, k, ctypes.byref(e), ctypes.byref(b),
                                                                 ctypes.byref(pp), ctypes.byref(idx), ctypes.byref(n)),
                       "get_pass_shape")
            out[name] = (e.value, b.value, idx.value, n.value, pp.value)
        return out

    def set_pass_shape(self, kind, index):
        """Force compiled shape `index` for pass `kind` (name or number)."""
        k = capi.PASS_KINDS.get(kind, kind)
        capi.check(self._h, capi.load().bfpsh_set_pass_shape(self._h, int(k), int(index)), "set_pass_shape")

    def _push_physics(self):
        rc = capi.load().bfpsh_set_physics(self._h, float(self.nu), capi.FORCING[self.forcing_type],
                                           float(self.famplitude), float(self.fk0), float(self.fk1),
                                           int(self.fmode))
        capi.check(self._h, rc, "set_physics")

    @property
    def iteration(self):
        it = ctypes.c_int()
        capi.check(self._h, capi.load().bfpsh_get_iteration(self._h, ctypes.byref(it)), "get_iteration")
        return it.value

    @iteration.setter
    def iteration(self, v):
        capi.check(self._h, capi.load().bfpsh_set_iteration(self._h, int(v)), "set_iteration")

    def cshape(self, nc=3):
        return (self.ny, self.nz, self.nx // 2 + 1, nc)

    # ---------------------------------------------------------------- fields
    def get_cfield(self, which=capi.CVORTICITY):
        out = np.zeros(self.cshape(), dtype=self.cdtype)
        capi.check(self._h, capi.load().bfpsh_get_cfield(self._h, which, out.ctypes.data), "get_cfield")
        return out

    @property
    def cvorticity(self):
        return self.get_cfield(capi.CVORTICITY)

    @cvorticity.setter
    def cvorticity(self, data):
        a = np.ascontiguousarray(np.asarray(data).reshape(self.cshape()), dtype=self.cdtype)
        capi.check(self._h, capi.load().bfpsh_put_cfield(self._h, capi.CVORTICITY, a.ctypes.data), "put_cfield")

    def get_current_fname(self):
        """<name>_checkpoint_<checkpoint>.h5 (vorticity_equation.hpp:98-104)."""
        return "%s_checkpoint_%d.h5" % (self.name, self.checkpoint)

    def update_checkpoint(self):
        """vorticity_equation::update_checkpoint (vorticity_equation.cpp:49-108):
        move to the next file once the current one holds checkpoints_per_file
        fields and not this iteration's; create the file (with the group
        vorticity/complex) when it does not exist."""
        fname = self.get_current_fname()
        if os.path.exists(fname):
            with h5io.File(fname) as f:
                stored = f.keys("vorticity/complex") if "vorticity/complex" in f else []
            if len(stored) >= self.checkpoints_per_file and str(self.iteration) not in stored:
                self.checkpoint += 1
        else:
            with h5io.File(fname, "w") as f:
                f.create_group("vorticity/complex")

    def io_checkpoint(self, read=True, data=None):
        """vorticity_equation::io_checkpoint (vorticity_equation.hpp:107-128).
        read=True: cvorticity <- vorticity/complex/<iteration> of the current
        checkpoint file (or `data`, an array in the same layout), then the
        reference's read projections (symmetrize, low_pass(kM),
        force_divfree).  read=False: update_checkpoint, then write cvorticity
        as vorticity/complex/<iteration> ({r, i} compound of rnumber)."""
        if read:
            if data is None:
                with h5io.File(self.get_current_fname()) as f:
                    data = f["vorticity/complex/%d" % self.iteration]
            self.cvorticity = data
            capi.check(self._h, capi.load().bfpsh_load_projection(self._h), "load_projection")
            return None
        self.update_checkpoint()
        with h5io.File(self.get_current_fname(), "a") as f:
            f["vorticity/complex/%d" % self.iteration] = self.cvorticity
        return None

    def compute_velocity(self):
        """cvelocity = Biot-Savart(cvorticity) + symmetrize (spectral)."""
        return self.get_cfield(capi.CVELOCITY)

    def impose_zero_modes(self):
        """vorticity_equation::impose_zero_modes (vorticity_equation.cpp:39-46):
        the k = 0 mode of cvelocity, cvorticity and the stage fields."""
        capi.check(self._h, capi.load().bfpsh_impose_zero_modes(self._h), "impose_zero_modes")

    def compute_vorticity(self, velocity=None):
        """vorticity_equation::compute_vorticity (vorticity_equation.cpp:171-204):
        cvorticity = i k x cvelocity for k^2 <= kM^2 (zero beyond), then
        symmetrize.  `velocity` (an array in the complex layout) first sets
        cvelocity; otherwise cvelocity is what the last call left there."""
        lib = capi.load()
        if velocity is not None:
            a = np.ascontiguousarray(np.asarray(velocity).reshape(self.cshape()), dtype=self.cdtype)
            capi.check(self._h, lib.bfpsh_put_cfield(self._h, capi.CVELOCITY, a.ctypes.data), "put_cfield")
        capi.check(self._h, lib.bfpsh_compute_vorticity(self._h), "compute_vorticity")

    def velocity_rspace(self):
        capi.check(self._h, capi.load().bfpsh_velocity_rspace(self._h), "velocity_rspace")

    def gradient_rspace(self):
        capi.check(self._h, capi.load().bfpsh_gradient_rspace(self._h), "gradient_rspace")

    def compute_pressure(self, pressure):
        """vorticity_equation::compute_pressure
        (bfps/cpp/vorticity_equation.cpp:441-520) of the velocity of
        cvorticity into a 1-component field (spectral)."""
        capi.check(self.handle, capi.load().bfpsh_compute_pressure(self.handle, pressure.id), "compute_pressure")
        pressure.real_space_representation = False

    def compute_Lagrangian_acceleration(self, acceleration):
        """vorticity_equation::compute_Lagrangian_acceleration
        (bfps/cpp/vorticity_equation.cpp:522-578) into a 3-component field
        (spectral)."""
        self._push_physics()
        capi.check(self.handle, capi.load().bfpsh_compute_lagrangian_acceleration(self.handle, acceleration.id),
                   "compute_lagrangian_acceleration")
        acceleration.real_space_representation = False

    def get_rfield(self, which=capi.RVELOCITY):
        nc = 3 if which == capi.RVELOCITY else 9
        out = np.zeros((self.nz, self.ny, self.nx, nc), dtype=self.rdtype)
        capi.check(self._h, capi.load().bfpsh_get_rfield(self._h, which, out.ctypes.data), "get_rfield")
        return out

    def cospectrum(self, a="velocity", b="velocity"):
        sel = {"vorticity": capi.CVORTICITY, "velocity": capi.CVELOCITY}
        out = np.zeros((self.kk.nshells, 3, 3))
        capi.check(self._h, capi.load().bfpsh_cospectrum(self._h, sel[a], sel[b], out.ctypes.data_as(capi._DP)),
                   "cospectrum")
        return out

    # ---------------------------------------------------------------- stepping
    def step(self, dt):
        self._push_physics()
        capi.check(self._h, capi.load().bfpsh_step(self._h, float(dt)), "step")

    def synchronize(self):
        capi.check(self._h, capi.load().bfpsh_synchronize(self._h), "synchronize")

    def time_steps(self, nsteps, dt, with_particles=False):
        self._push_physics()
        ms = ctypes.c_double()
        capi.check(self._h, capi.load().bfpsh_time_steps(self._h, int(nsteps), float(dt), int(bool(with_particles)),
                                                         ctypes.byref(ms)), "time_steps")
        return ms.value

    def profile(self, enable=True):
        capi.check(self._h, capi.load().bfpsh_profile_enable(self._h, int(enable)), "profile_enable")

    def profile_report(self):
        out = {}
        lib = capi.load()
        i = 0
        buf = ctypes.create_string_buffer(128)
        while True:
            ms = ctypes.c_double()
            cnt = ctypes.c_int()
            if lib.bfpsh_profile_get(self._h, i, buf, 128, ctypes.byref(ms), ctypes.byref(cnt)) != 0:
                break
            out[buf.value.decode()] = (ms.value, cnt.value)
            i += 1
        return out

    def device_bytes(self):
        b = ctypes.c_size_t()
        capi.check(self._h, capi.load().bfpsh_device_bytes(self._h, ctypes.byref(b)), "device_bytes")
        return b.value


def compute_gradient(kk, src, dst):
    """The free function compute_gradient(kk, src, dst)
    (bfps/cpp/field.cpp:1042-1101) on standalone fields of one context:
    src ONE -> dst THREE or THREE -> THREExTHREE, dst component
    j ncomp(src) + i = i k_j src_i for k^2 < kM^2, zero elsewhere.  `kk` is
    the context's kspace (accepted for the reference's signature)."""
    del kk
    if src.fs is not dst.fs:
        raise capi.BfpshError("compute_gradient needs two fields of one context")
    capi.check(src.fs.handle, capi.load().bfpsh_field_compute_gradient(src.fs.handle, src.id, dst.id),
               "compute_gradient")
    dst.real_space_representation = False


class field:
    """Device-resident ``field<rnumber, FFTW, fc>`` of a context
    (bfps/cpp/field.hpp:47-200): ncomp = 1 (ONE), 3 (THREE) or 9
    (THREExTHREE).  Transforms are the reference's unnormalised in-place
    ``dft`` (r2c) / ``ift`` (c2r); ``real_space_representation`` tracks which
    one the data is in, as the reference's member of that name."""

    _FC = {"ONE": 1, "THREE": 3, "THREExTHREE": 9, 1: 1, 3: 3, 9: 9}

    def __init__(self, fs, fc="THREE"):
        self.fs = fs
        # anything else reaches the C ABI, which rejects it
        self.ncomp = self._FC.get(fc, fc if isinstance(fc, int) else 0)
        self._id = ctypes.c_int(-1)
        capi.check(fs.handle, capi.load().bfpsh_field_create(fs.handle, self.ncomp, ctypes.byref(self._id)),
                   "field_create")
        self.real_space_representation = False

    @property
    def id(self):
        return self._id.value

    def close(self):
        if self._id.value >= 0 and self.fs.handle:
            capi.check(self.fs.handle, capi.load().bfpsh_field_destroy(self.fs.handle, self._id.value),
                       "field_destroy")
        self._id.value = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rshape(self):
        return (self.fs.nz, self.fs.ny, self.fs.nx, self.ncomp)

    def cshape(self):
        return self.fs.cshape(self.ncomp)

    def put(self, data, real_space):
        """Host data in the reference layout ([nz][ny][nx][nc] real or
        [ny][nz][nx/2+1][nc] complex)."""
        if real_space:
            a = np.ascontiguousarray(np.asarray(data).reshape(self.rshape()), dtype=self.fs.rdtype)
        else:
            a = np.ascontiguousarray(np.asarray(data).reshape(self.cshape()), dtype=self.fs.cdtype)
        capi.check(self.fs.handle, capi.load().bfpsh_field_put(self.fs.handle, self.id, int(bool(real_space)),
                                                               a.ctypes.data), "field_put")
        self.real_space_representation = bool(real_space)

    def get(self):
        if self.real_space_representation:
            out = np.zeros(self.rshape(), dtype=self.fs.rdtype)
        else:
            out = np.zeros(self.cshape(), dtype=self.fs.cdtype)
        capi.check(self.fs.handle, capi.load().bfpsh_field_get(self.fs.handle, self.id,
                                                               int(self.real_space_representation),
                                                               out.ctypes.data), "field_get")
        return out

    def get_rdata(self):
        assert self.real_space_representation
        return self.get()

    def get_cdata(self):
        assert not self.real_space_representation
        return self.get()

    def dft(self):
        """field::dft (bfps/cpp/field.cpp:154-162)."""
        capi.check(self.fs.handle, capi.load().bfpsh_field_dft(self.fs.handle, self.id), "field_dft")
        self.real_space_representation = False

    def ift(self):
        """field::ift (bfps/cpp/field.cpp:144-152)."""
        capi.check(self.fs.handle, capi.load().bfpsh_field_ift(self.fs.handle, self.id), "field_ift")
        self.real_space_representation = True

    def symmetrize(self):
        """field::symmetrize (bfps/cpp/field.cpp:900-978)."""
        capi.check(self.fs.handle, capi.load().bfpsh_field_symmetrize(self.fs.handle, self.id), "field_symmetrize")

    def copy_from(self, which="vorticity"):
        """Spectral copy of the solver's cvorticity or its Biot-Savart velocity
        (bfps/cpp/full_code/NSVE.cpp:115,124)."""
        sel = {"vorticity": capi.CVORTICITY, "velocity": capi.CVELOCITY}[which]
        capi.check(self.fs.handle, capi.load().bfpsh_field_copy_from(self.fs.handle, self.id, sel),
                   "field_copy_from")
        self.real_space_representation = False

    def normalize(self):
        """field::normalize (bfps/cpp/field.cpp:891-898): data /= npoints."""
        capi.check(self.fs.handle, capi.load().bfpsh_field_normalize(self.fs.handle, self.id), "field_normalize")

    def compute_rspace_stats(self, max_estimate, nbins=256):
        """field::compute_rspace_stats (bfps/cpp/field.cpp:661-850) without the
        HDF5 writes: (moments [10][nvals], histograms [nbins][nvals]) with
        nvals = 4 for a vector field (the magnitude last), else ncomp.
        max_estimate holds nvals values (the reference's max_estimate_vector)."""
        nvals = 4 if self.ncomp == 3 else self.ncomp
        me = np.ascontiguousarray(max_estimate, dtype=np.float64).reshape(nvals)
        mom = np.zeros((10, nvals))
        hist = np.zeros((nbins, nvals), dtype=np.int64)
        capi.check(self.fs.handle, capi.load().bfpsh_field_rspace_stats(
            self.fs.handle, self.id, int(nbins), me.ctypes.data_as(capi._DP), mom.ctypes.data_as(capi._DP),
            hist.ctypes.data_as(capi._LP)), "field_rspace_stats")
        return mom, hist

    def compute_stats(self, max_estimate, nbins=256):
        """field::compute_stats (bfps/cpp/field.cpp:980-1035) of a vector
        field, returning what the reference writes: {"moments", "histograms",
        "cospectrum"}.  max_estimate is the scalar of the reference's call;
        the magnitude's range is sqrt(3) times it.  A real-space field is
        left spectral and normalised, a spectral one real, as the reference
        leaves them."""
        assert self.ncomp == 3
        me = [max_estimate] * 3 + [max_estimate * np.sqrt(3.0)]
        out = {}
        if self.real_space_representation:
            out["moments"], out["histograms"] = self.compute_rspace_stats(me, nbins)
            self.dft()
            self.normalize()
            out["cospectrum"] = self.cospectrum(self)
        else:
            out["cospectrum"] = self.cospectrum(self)
            self.ift()
            out["moments"], out["histograms"] = self.compute_rspace_stats(me, nbins)
        return out

    def cospectrum(self, other):
        """kspace::cospectrum of two spectral 3-component fields, [shell][i][j]."""
        out = np.zeros((self.fs.kk.nshells, 3, 3))
        capi.check(self.fs.handle, capi.load().bfpsh_field_cospectrum(self.fs.handle, self.id, other.id,
                                                                      out.ctypes.data_as(capi._DP)),
                   "field_cospectrum")
        return out


class particles_system:
    """Tracers of one context (particles_system on one rank).

    neighbours n in [1, 10], smoothness m in [0, 2]
    (particles_generic_interp<double, n, m>), integration_steps = AB order
    (1..6), step_idx as passed by NSVEparticles (iteration + 1)."""

    def __init__(self, fs, nparticles, neighbours, smoothness, integration_steps=4, step_idx=1):
        self.fs = fs
        self.np = int(nparticles)
        self.nsteps = int(integration_steps)
        rc = capi.load().bfpsh_particles_create(fs.handle, self.np, int(neighbours), int(smoothness),
                                                self.nsteps, int(step_idx))
        capi.check(fs.handle, rc, "particles_create")
        fs.particles = self

    def put(self, pos, rhs=None, idx=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.np, 3)
        r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).reshape(self.nsteps, self.np, 3)
        i = None if idx is None else np.ascontiguousarray(idx, dtype=np.int64)
        rc = capi.load().bfpsh_particles_put(self.fs.handle, pos.ctypes.data_as(capi._DP),
                                             None if r is None else r.ctypes.data_as(capi._DP),
                                             None if i is None else i.ctypes.data_as(capi._LP))
        capi.check(self.fs.handle, rc, "particles_put")

    def get(self):
        pos = np.zeros((self.np, 3))
        rhs = np.zeros((self.nsteps, self.np, 3))
        idx = np.zeros(self.np, dtype=np.int64)
        si = ctypes.c_int()
        rc = capi.load().bfpsh_particles_get(self.fs.handle, pos.ctypes.data_as(capi._DP),
                                             rhs.ctypes.data_as(capi._DP), idx.ctypes.data_as(capi._LP),
                                             ctypes.byref(si))
        capi.check(self.fs.handle, rc, "particles_get")
        return pos, rhs, idx, si.value

    def getParticlesPositions(self):
        return self.get()[0]

    def getParticlesRhs(self):
        return self.get()[1]

    def getParticlesIndexes(self):
        return self.get()[2]

    def get_step_idx(self):
        return self.get()[3]

    def completeLoop(self, dt):
        capi.check(self.fs.handle, capi.load().bfpsh_particles_complete_loop(self.fs.handle, float(dt)),
                   "particles_complete_loop")

    def sample(self, which="velocity"):
        sel = capi.RVELOCITY if which == "velocity" else capi.RGRADIENT
        nc = 3 if which == "velocity" else 9
        out = np.zeros((self.np, nc))
        rc = capi.load().bfpsh_particles_sample(self.fs.handle, sel, out.ctypes.data_as(capi._DP))
        capi.check(self.fs.handle, rc, "particles_sample")
        return out

    def write_checkpoint(self, fname, iteration):
        """The tracer part of a checkpoint (particles_output_hdf5,
        bfps/cpp/particles/particles_output_hdf5.hpp:110-300):
        tracers0/state/<iteration> [np][3] and tracers0/rhs/<iteration>
        [nsteps][np][3], FP64, in global-index order."""
        pos, rhs, _, _ = self.get()
        with h5io.File(fname, "a") as f:
            f["tracers0/state/%d" % iteration] = pos
            f["tracers0/rhs/%d" % iteration] = rhs

    def read_checkpoint(self, fname, iteration):
        """Loads state and history as particles_system_builder does
        (particles_system_builder.hpp:193-233): tracers0/state/<iteration>
        and tracers0/rhs/<iteration>, global indexes 0..np-1 in file order."""
        with h5io.File(fname) as f:
            pos = f["tracers0/state/%d" % iteration]
            rhs = f["tracers0/rhs/%d" % iteration]
        if pos.shape != (self.np, 3) or rhs.shape != (self.nsteps, self.np, 3):
            raise capi.BfpshError("checkpoint holds %s / %s, the system has %d tracers and %d steps of history"
                                  % (pos.shape, rhs.shape, self.np, self.nsteps))
        self.put(pos, rhs)

    def sample_compute_field(self, f, out=None):
        """sample_compute_field of a standalone field in real space
        (bfps/cpp/particles/particles_system.hpp:143-170): [np][ncomp]."""
        if out is None:
            out = np.zeros((self.np, f.ncomp))
        rc = capi.load().bfpsh_particles_sample_field(self.fs.handle, f.id, out.ctypes.data_as(capi._DP))
        capi.check(self.fs.handle, rc, "particles_sample_field")
        return out


def NSVE_do_stats(fs, max_velocity_estimate=1.0, max_vorticity_estimate=1.0, nbins=256):
    """NSVE::do_stats (bfps/cpp/full_code/NSVE.cpp:101-135) without the
    statistics file: field::compute_stats of the vorticity and of the
    velocity, returned as {"vorticity": ..., "velocity": ...} (moments,
    histograms, cospectrum).  Also returns the work field, which is left
    holding the real-space velocity as in the reference (NSVEparticles'
    do_stats samples it)."""
    tmp = field(fs, "THREE")
    out = {}
    tmp.copy_from("vorticity")
    out["vorticity"] = tmp.compute_stats(max_vorticity_estimate / np.sqrt(3.0), nbins)
    tmp.copy_from("velocity")
    out["velocity"] = tmp.compute_stats(max_velocity_estimate / np.sqrt(3.0), nbins)
    return out, tmp


def NSVEparticles_do_stats(fs, ps, max_velocity_estimate=1.0, max_vorticity_estimate=1.0, nbins=256):
    """NSVEparticles::do_stats (bfps/cpp/full_code/NSVEparticles.cpp:70-98)
    without the files: the fluid statistics, then the velocity and the
    Lagrangian acceleration sampled at the tracers ([np][3] each, in
    global-index order)."""
    out, tmp = NSVE_do_stats(fs, max_velocity_estimate, max_vorticity_estimate, nbins)
    out["tracers0/velocity"] = ps.sample_compute_field(tmp)
    fs.compute_Lagrangian_acceleration(tmp)
    tmp.ift()
    out["tracers0/acceleration"] = ps.sample_compute_field(tmp)
    tmp.close()
    return out


def NSVEparticles_write_checkpoint(fs, ps):
    """NSVEparticles::write_checkpoint (NSVEparticles.cpp:44-56): the
    vorticity, then the tracers into the same checkpoint file."""
    fs.io_checkpoint(False)
    if ps is not None:
        ps.write_checkpoint(fs.get_current_fname(), fs.iteration)


def NSVEparticles_read_checkpoint(fs, ps):
    """NSVE::initialize + NSVEparticles::initialize (NSVE.cpp:35-60,
    NSVEparticles.cpp:7-30): vorticity and tracers of fs.iteration from the
    current checkpoint file."""
    fs.io_checkpoint(True)
    if ps is not None:
        ps.read_checkpoint(fs.get_current_fname(), fs.iteration)


def NSVEparticles_step(fs, dt):
    """NSVEparticles::step: tracers advance with u(t), then the fluid step."""
    fs._push_physics()
    capi.check(fs.handle, capi.load().bfpsh_nsve_particles_step(fs.handle, float(dt)), "nsve_particles_step")


def wisdom_import(path):
    """Load pass wisdom from a text file (fftwf_import_wisdom_from_filename,
    bfps/cpp/full_code/main_code.hpp:101-104); a missing file is no error."""
    rc = capi.load().bfpsh_wisdom_import(os.fsencode(path))
    if rc != 0:
        raise capi.BfpshError("wisdom_import failed (%d): %s" % (rc, path))


def wisdom_export(path):
    """Save the pass wisdom (fftwf_export_wisdom_to_filename, main_code.hpp:145-149)."""
    rc = capi.load().bfpsh_wisdom_export(os.fsencode(path))
    if rc != 0:
        raise capi.BfpshError("wisdom_export failed (%d): %s" % (rc, path))


def wisdom_forget():
    capi.load().bfpsh_wisdom_forget()


class ensemble:
    """Independent boxes of sizes ``sizes`` (box 2 pi, one precision) advanced
    together.  ``boxes[i]`` is a ``vorticity_equation`` over box i's context:
    set its physics, fields and tracers as for a single box; ``step`` then
    runs one NSVEparticles step of every box (box i with ``dt[i]``)."""

    def __init__(self, sizes, precision="single", dealias="SMOOTH", device=0):
        lib = capi.load()
        sizes = [int(n) for n in sizes]
        h = ctypes.c_void_p()
        arr = (ctypes.c_int * len(sizes))(*sizes)
        rc = lib.bfpsh_ensemble_create(ctypes.byref(h), len(sizes), arr,
                                       capi.FP32 if precision == "single" else capi.FP64,
                                       capi.SMOOTH if dealias == "SMOOTH" else capi.TWO_THIRDS, device)
        if rc != 0:
            msg = lib.bfpsh_ensemble_last_error(h)
            lib.bfpsh_ensemble_destroy(h)
            raise capi.BfpshError("bfpsh_ensemble_create failed (%d): %s" % (rc, msg.decode() if msg else ""))
        self._h = h
        self.sizes = sizes
        self.boxes = []
        for i, n in enumerate(sizes):
            b = ctypes.c_void_p()
            self._check(lib.bfpsh_ensemble_box(h, i, ctypes.byref(b)), "ensemble_box")
            self.boxes.append(vorticity_equation("box%d" % i, n, precision=precision, dealias=dealias,
                                                 _handle=b.value))

    def _check(self, rc, what):
        if rc != 0:
            msg = capi.load().bfpsh_ensemble_last_error(self._h)
            raise capi.BfpshError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def _dt(self, dt):
        d = np.broadcast_to(np.asarray(dt, dtype=np.float64), (len(self.sizes),))
        return np.ascontiguousarray(d)

    def step(self, dt, with_particles=True):
        for b in self.boxes:
            b._push_physics()
        d = self._dt(dt)
        self._check(capi.load().bfpsh_ensemble_step(self._h, d.ctypes.data_as(capi._DP), int(bool(with_particles))),
                    "ensemble_step")

    def time_steps(self, nsteps, dt, with_particles=True):
        for b in self.boxes:
            b._push_physics()
        d = self._dt(dt)
        ms = ctypes.c_double()
        self._check(capi.load().bfpsh_ensemble_time_steps(self._h, int(nsteps), d.ctypes.data_as(capi._DP),
                                                          int(bool(with_particles)), ctypes.byref(ms)),
                    "ensemble_time_steps")
        return ms.value

    def step_separate(self, dt, with_particles=True):
        """The same step issued box by box (no batching, no graph)."""
        for b in self.boxes:
            b._push_physics()
        d = self._dt(dt)
        self._check(capi.load().bfpsh_ensemble_step_separate(self._h, d.ctypes.data_as(capi._DP),
                                                             int(bool(with_particles))), "ensemble_step_separate")

    def time_steps_separate(self, nsteps, dt, with_particles=True):
        for b in self.boxes:
            b._push_physics()
        d = self._dt(dt)
        ms = ctypes.c_double()
        self._check(capi.load().bfpsh_ensemble_time_steps_separate(
            self._h, int(nsteps), d.ctypes.data_as(capi._DP), int(bool(with_particles)), ctypes.byref(ms)),
            "ensemble_time_steps_separate")
        return ms.value

    def synchronize(self):
        self._check(capi.load().bfpsh_ensemble_synchronize(self._h), "ensemble_synchronize")

    def close(self):
        if getattr(self, "_h", None):
            for b in self.boxes:
                b._h = None
            capi.load().bfpsh_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
