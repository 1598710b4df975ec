/* bfpsh -- C ABI of the MI355X (gfx950) backend for bfps's vorticity-equation
 * time step and tracer loop.
 *
 * One context holds one periodic box entirely in device memory (cubic N^3,
 * N in {16, 32, 48, 64, 128, 256, 512, 1024}) plus an optional tracer set.
 * All calls are ordered on the context's HIP stream and are not re-entrant
 * per context; every function returns 0 on success and a non-zero
 * BFPSH_E* code otherwise, with a message in bfpsh_last_error().
 * No exceptions cross this boundary.
 *
 * Host-side array layouts are the reference's (bfps/cpp/field.cpp:85-101):
 *   complex field: rnumber {re, im} [ny][nz][nx/2+1][ncomp]  (FFTW transposed)
 *   real field:    rnumber          [nz][ny][nx][ncomp]      (padding dropped)
 *   tracers:       double positions [np][3], history [nsteps][np][3]
 * where rnumber is float (BFPSH_FP32) or double (BFPSH_FP64), as
 * vorticity_equation<rnumber, FFTW>.
 */
#ifndef BFPSH_H
#define BFPSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfpsh_ctx bfpsh_ctx;

/* precision: field storage type (rnumber of the reference) */
enum { BFPSH_FP32 = 0, BFPSH_FP64 = 1 };
/* dealiasing: kspace_dealias_type (bfps/cpp/kspace.hpp:40) */
enum { BFPSH_TWO_THIRDS = 0, BFPSH_SMOOTH = 1 };
/* forcing_type strings of vorticity_equation::add_forcing
 * (bfps/cpp/vorticity_equation.cpp:241-288), plus a fixed energy-injection
 * rate (no reference counterpart): the linear band forcing with amplitude
 * famplitude / (2 E_band), E_band the kinetic energy of the band
 * fk0 <= |k| <= fk1 of the stage's source field, so that the band gains
 * energy at the rate famplitude. */
enum { BFPSH_FORCING_NONE = 0, BFPSH_FORCING_LINEAR = 1, BFPSH_FORCING_KOLMOGOROV = 2,
       BFPSH_FORCING_FIXED_ENERGY = 3 };
/* field selectors */
enum {
    BFPSH_CVORTICITY = 0, /* vorticity_equation::cvorticity (spectral) */
    BFPSH_CVELOCITY = 1,  /* Biot-Savart of cvorticity (compute_velocity) */
    BFPSH_RVELOCITY = 5,  /* real-space velocity (3 comps) */
    BFPSH_RGRADIENT = 6   /* real-space velocity gradient (9 comps, j*3+i = d_j u_i) */
};
/* error codes */
enum {
    BFPSH_OK = 0,
    BFPSH_EINVAL = 1,  /* bad argument or unsupported size */
    BFPSH_EHIP = 2,    /* HIP runtime error */
    BFPSH_ESTATE = 3,  /* call not valid in the current state */
    BFPSH_ENOMEM = 4
};

/* Replaces the constructor vorticity_equation<rnumber, FFTW>(name, nx, ny, nz,
 * dkx, dky, dkz, plan_rigor) (bfps/cpp/vorticity_equation.cpp:110-156) and
 * the kspace<FFTW, SMOOTH> it builds (bfps/cpp/kspace.cpp:34-167).  The
 * reference hard-codes SMOOTH; `dealias` also allows TWO_THIRDS.
 * device: HIP device ordinal.  All fields start at zero. */
int bfpsh_create(bfpsh_ctx** out, int nx, int ny, int nz, double dkx, double dky, double dkz,
                 int precision, int dealias, int device);
/* Plan rigor of the transforms: replaces the `plan_rigor` argument of the
 * field / vorticity_equation constructors, which picks FFTW's plan search
 * (FFTW_ESTIMATE / FFTW_MEASURE / FFTW_PATIENT; bfps/cpp/field.cpp:105-118,
 * bfps/cpp/fftw_interface.hpp:30-35).  Here it picks how the tile shape of
 * each strided FFT pass kind is chosen among the compiled shapes:
 * ESTIMATE takes the shape the pass wisdom holds, else the built-in one;
 * MEASURE and PATIENT take the wisdom's shape when a search of at least
 * their rigor chose it, else time every compiled shape (3 launches each on
 * one scratch plane for MEASURE, 10 for PATIENT) and keep the fastest in
 * the wisdom.  Boxes below 256^3 have one shape.  At MEASURE and PATIENT
 * the first tracer loop or sample of a tracer set also times the velocity
 * loop of each applicable tracer kernel form (column walk, bricks, global
 * gather) on a device copy of the tracer state, which it allocates for the
 * time it runs, and keeps the fastest when it is more than 10 % faster
 * than the first applicable one (BFPSH_TRACER_KERNEL=column|brick|gather
 * in the environment forces a form, BFPSH_TRACER_VERBOSE prints the
 * trials). */
enum { BFPSH_ESTIMATE = 0, BFPSH_MEASURE = 1, BFPSH_PATIENT = 2 };
/* strided pass kinds: inverse and forward transforms along kz (z) and ky
 * (y), the vorticity step's fused first pass (Biot-Savart + inverse z) and
 * its fused last pass (forward y + dealias, curl, forcing, projection and
 * the Runge-Kutta update) */
enum { BFPSH_PASS_Z_INV = 0, BFPSH_PASS_Y_INV = 1, BFPSH_PASS_Z_FWD = 2, BFPSH_PASS_Y_FWD = 3,
       BFPSH_PASS_BS = 4, BFPSH_PASS_Y3 = 5, BFPSH_PASS_KINDS = 6 };

/* Replaces ~vorticity_equation (bfps/cpp/vorticity_equation.cpp:158-169). */
int bfpsh_destroy(bfpsh_ctx* ctx);
const char* bfpsh_last_error(const bfpsh_ctx* ctx);

/* Sets the public physics members that NSVE::initialize writes
 * (bfps/cpp/full_code/NSVE.cpp:46-52): nu, forcing_type, famplitude, fk0,
 * fk1, fmode. */
int bfpsh_set_physics(bfpsh_ctx* ctx, double nu, int forcing, double famplitude, double fk0,
                      double fk1, int fmode);
/* kspace metadata (bfps/cpp/kspace.cpp:97-166): out[0] = kM, out[1] = kM2,
 * out[2] = dk, out[3] = nshells; kshell/nshell (may be NULL) sized nshells. */
int bfpsh_kspace_info(bfpsh_ctx* ctx, double* out4, double* kshell, int64_t* nshell);

/* Replaces `*cvorticity = host_data` / field::io read of vorticity/complex
 * (bfps/cpp/field.hpp:193-200, field.cpp:336-341) without the projection;
 * which = BFPSH_CVELOCITY sets `*cvelocity` (the `u` member the
 * vorticity_equation reads in compute_vorticity). */
int bfpsh_put_cfield(bfpsh_ctx* ctx, int which, const void* host);
/* Replaces reading field::get_cdata() of cvorticity / cvelocity
 * (bfps/cpp/field.hpp:140-146).  The kx = nx/2 plane is returned as zero. */
int bfpsh_get_cfield(bfpsh_ctx* ctx, int which, void* host);
/* Real-space field (BFPSH_RVELOCITY or BFPSH_RGRADIENT) as computed by the
 * last bfpsh_velocity_rspace / bfpsh_gradient_rspace. */
int bfpsh_get_rfield(bfpsh_ctx* ctx, int which, void* host);

/* Replaces the projections of vorticity_equation::io_checkpoint(read=true)
 * (bfps/cpp/vorticity_equation.hpp:107-128): symmetrize (field.cpp:340),
 * low_pass(kM), force_divfree. */
int bfpsh_load_projection(bfpsh_ctx* ctx);
/* Replaces vorticity_equation::compute_vorticity()
 * (bfps/cpp/vorticity_equation.cpp:171-204): cvorticity = i k x cvelocity
 * for k^2 <= kM^2, zero beyond, then symmetrize.  cvelocity is what the last
 * bfpsh_put_cfield(BFPSH_CVELOCITY) set (a step or a velocity computation
 * in between overwrites it, as they overwrite the reference's `u`). */
int bfpsh_compute_vorticity(bfpsh_ctx* ctx);
/* Replaces vorticity_equation::impose_zero_modes()
 * (bfps/cpp/vorticity_equation.cpp:39-46): the k = 0 mode of cvelocity,
 * cvorticity and the two stage fields set to zero. */
int bfpsh_impose_zero_modes(bfpsh_ctx* ctx);
/* Replaces vorticity_equation::step(dt) (bfps/cpp/vorticity_equation.cpp:358-439). */
int bfpsh_step(bfpsh_ctx* ctx, double dt);
/* Replaces `compute_velocity(cvorticity); cvelocity->ift();`
 * (bfps/cpp/full_code/NSVEparticles.cpp:36-37). */
int bfpsh_velocity_rspace(bfpsh_ctx* ctx);
/* Replaces compute_gradient(kk, cvelocity, grad) followed by grad->ift()
 * (bfps/cpp/field.cpp:1042-1101) on the velocity of cvorticity. */
int bfpsh_gradient_rspace(bfpsh_ctx* ctx);
/* Replaces kspace::cospectrum (bfps/cpp/kspace.cpp:521-598) of two 3-comp
 * spectral fields a, b (BFPSH_CVORTICITY / BFPSH_CVELOCITY); out sized
 * nshells*9 doubles, [shell][i][j]. */
int bfpsh_cospectrum(bfpsh_ctx* ctx, int a, int b, double* out);

/* Standalone device field of ncomp (1, 3 or 9) components with the
 * reference's in-place transforms: replaces the construction of
 * field<rnumber, FFTW, fc> (bfps/cpp/field.cpp:37-121).  *id receives the
 * handle; the field starts at zero in the spectral representation.  Modes
 * kx = nx/2 are not represented (they lie outside every dealiasing sphere;
 * put ignores them, get returns zeros). */
int bfpsh_field_create(bfpsh_ctx* ctx, int ncomp, int* id);
/* Replaces ~field (bfps/cpp/field.cpp:123-142). */
int bfpsh_field_destroy(bfpsh_ctx* ctx, int id);
/* Replaces field::ift (bfps/cpp/field.cpp:144-152): unnormalised c2r in place. */
int bfpsh_field_ift(bfpsh_ctx* ctx, int id);
/* Replaces field::dft (bfps/cpp/field.cpp:154-162): unnormalised r2c in place. */
int bfpsh_field_dft(bfpsh_ctx* ctx, int id);
/* Replaces field::symmetrize (bfps/cpp/field.cpp:900-978). */
int bfpsh_field_symmetrize(bfpsh_ctx* ctx, int id);
/* Host copies in the reference layouts: real_space != 0 -> rnumber
 * [nz][ny][nx][ncomp]; 0 -> rnumber {re, im} [ny][nz][nx/2+1][ncomp]
 * (field::get_rdata / get_cdata, bfps/cpp/field.hpp:135-146). */
int bfpsh_field_put(bfpsh_ctx* ctx, int id, int real_space, const void* host);
/* Replaces the free function compute_gradient(kk, src, dst)
 * (bfps/cpp/field.cpp:1042-1101) on standalone fields: src ONE -> dst THREE
 * or src THREE -> dst THREExTHREE (component j ncomp(src) + i = i k_j src_i
 * for k^2 < kM^2, zero elsewhere); src spectral, dst left spectral. */
int bfpsh_field_compute_gradient(bfpsh_ctx* ctx, int src_id, int dst_id);
int bfpsh_field_get(bfpsh_ctx* ctx, int id, int real_space, void* host);
/* Replaces `*tmp_vec_field = fs->cvorticity->get_cdata()` and the same for
 * cvelocity (bfps/cpp/full_code/NSVE.cpp:115,124): which = BFPSH_CVORTICITY
 * or BFPSH_CVELOCITY into a 3-component field. */
int bfpsh_field_copy_from(bfpsh_ctx* ctx, int id, int which);
/* kspace::cospectrum (bfps/cpp/kspace.cpp:521-598) of two 3-component
 * standalone fields (spectral); out nshells*9 doubles. */
int bfpsh_field_cospectrum(bfpsh_ctx* ctx, int id_a, int id_b, double* out);
/* No counterpart in the reference: bfpsh_field_cospectrum without any
 * allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its shell sums use a fixed device array of at most 1024 shells (73.7 KB,
 * one per GPU, held until the process ends) instead of a per-call buffer;
 * more shells are rejected (BFPSH_EINVAL).  Same arguments and output as
 * bfpsh_field_cospectrum.  Calls that overlap on one GPU share the array
 * and race. */
int bfpsh_field_cospectrum_noalloc(bfpsh_ctx* ctx, int id_a, int id_b, double* out);
/* Replaces field::normalize (bfps/cpp/field.cpp:891-898): every value /=
 * nx*ny*nz, in the field precision. */
int bfpsh_field_normalize(bfpsh_ctx* ctx, int id);
/* Replaces field::compute_rspace_stats (bfps/cpp/field.cpp:661-850) of a
 * field in real space, without the HDF5 writes: moments [10][nvals] (row 0
 * min, row 9 max, both starting at 0 -- the magnitude's min at its estimate
 * --, rows 1..8 the means of value^1..value^8) and hist [nbins][nvals]
 * counts; nvals = 4 for a 3-component field (value 3 = magnitude), else
 * ncomp; max_estimate[nvals] sets the histogram ranges (-M..M, the
 * magnitude's 0..M).  1 <= nbins <= 4096. */
int bfpsh_field_rspace_stats(bfpsh_ctx* ctx, int id, int nbins, const double* max_estimate, double* moments,
                             int64_t* hist);
/* Replaces vorticity_equation::compute_pressure
 * (bfps/cpp/vorticity_equation.cpp:441-520) into a 1-component field
 * (spectral); the real-space velocity it takes is made here from
 * cvorticity. */
int bfpsh_compute_pressure(bfpsh_ctx* ctx, int id);
/* Replaces vorticity_equation::compute_Lagrangian_acceleration
 * (bfps/cpp/vorticity_equation.cpp:522-578) into a 3-component field
 * (spectral). */
int bfpsh_compute_lagrangian_acceleration(bfpsh_ctx* ctx, int id);

/* Replaces particles_system_builder(...) (bfps/cpp/particles/particles_system_builder.hpp:238-257)
 * for one rank: np tracers, interpolation neighbours n in [1,10] and
 * smoothness m in [0,2] (particles_generic_interp.hpp:10-313), nsteps
 * Adams-Bashforth history length in [1,6], initial step_idx. */
int bfpsh_particles_create(bfpsh_ctx* ctx, int64_t np, int n, int m, int nsteps, int step_idx);
/* State as read by particles_input_hdf5 (positions [np][3], history
 * [nsteps][np][3], global indexes [np]; idx may be NULL for 0..np-1). */
int bfpsh_particles_put(bfpsh_ctx* ctx, const double* pos, const double* rhs, const int64_t* idx);
/* State as written by particles_output_hdf5::save: sorted by global index;
 * rhs may be NULL; step_idx may be NULL. */
int bfpsh_particles_get(bfpsh_ctx* ctx, double* pos, double* rhs, int64_t* idx, int* step_idx);
/* Replaces particles_system::completeLoop(dt) (bfps/cpp/particles/particles_system.hpp:210-217)
 * using the velocity of the current cvorticity. */
int bfpsh_particles_complete_loop(bfpsh_ctx* ctx, double dt);
/* Replaces particles_system::sample_compute_field for the real-space
 * velocity (3 comps, BFPSH_RVELOCITY) or gradient (9 comps, BFPSH_RGRADIENT);
 * out [np][ncomp], in the current (index-sorted) order. */
int bfpsh_particles_sample(bfpsh_ctx* ctx, int which, double* out);
/* Replaces particles_system::sample_compute_field(field<rnumber, FFTW, fc>)
 * (bfps/cpp/particles/particles_system.hpp:143-170) for a standalone field
 * in real space: out [np][ncomp] in global-index order.  As in the
 * reference, a one-component FP32 field is not sampled (out is left
 * untouched). */
int bfpsh_particles_sample_field(bfpsh_ctx* ctx, int id, double* out);

/* Replaces NSVEparticles::step (bfps/cpp/full_code/NSVEparticles.cpp:33-41):
 * velocity of omega(t) -> tracers completeLoop(dt) -> vorticity step(dt),
 * with the tracer velocity taken from the first RK stage (same data). */
int bfpsh_nsve_particles_step(bfpsh_ctx* ctx, double dt);

/* Ensembles: many independent boxes (sizes[i]^3, one precision, box 2 pi,
 * per-box physics and tracers) advanced together.  Each NSVEparticles::step
 * (bfps/cpp/full_code/NSVEparticles.cpp:33-41) of every box runs as one HIP
 * graph launch of batched field kernels (one launch per kernel and box size)
 * followed by the tracer loop of each box.  The reference runs one box per
 * job; the box contexts returned by bfpsh_ensemble_box accept every bfpsh_*
 * call above (physics, fields, tracers) and are owned by the ensemble. */
typedef struct bfpsh_ensemble bfpsh_ensemble;
int bfpsh_ensemble_create(bfpsh_ensemble** out, int nboxes, const int* sizes, int precision, int dealias,
                          int device);
int bfpsh_ensemble_destroy(bfpsh_ensemble* e);
const char* bfpsh_ensemble_last_error(const bfpsh_ensemble* e);
int bfpsh_ensemble_box(bfpsh_ensemble* e, int i, bfpsh_ctx** box);
/* One step of every box, box i with time step dt[i]. */
int bfpsh_ensemble_step(bfpsh_ensemble* e, const double* dt, int with_particles);
int bfpsh_ensemble_time_steps(bfpsh_ensemble* e, int nsteps, const double* dt, int with_particles, double* ms);
/* The same step issued box by box (one launch sequence per box, no batching
 * or graph): the per-box baseline the batched form is measured against. */
int bfpsh_ensemble_step_separate(bfpsh_ensemble* e, const double* dt, int with_particles);
int bfpsh_ensemble_time_steps_separate(bfpsh_ensemble* e, int nsteps, const double* dt, int with_particles,
                                       double* ms);
int bfpsh_ensemble_synchronize(bfpsh_ensemble* e);

int bfpsh_synchronize(bfpsh_ctx* ctx);
/* Current iteration counter (vorticity_equation::iteration). */
int bfpsh_get_iteration(bfpsh_ctx* ctx, int* it);
int bfpsh_set_iteration(bfpsh_ctx* ctx, int it);

/* Measurement: run `nsteps` steps (with tracers when with_particles != 0)
 * between two HIP events on the context stream; *ms = elapsed time. */
int bfpsh_time_steps(bfpsh_ctx* ctx, int nsteps, double dt, int with_particles, double* ms);
/* Per-kernel timing with HIP events around every launch (adds overhead;
 * for measurement runs only).  enable != 0 resets the counters. */
int bfpsh_profile_enable(bfpsh_ctx* ctx, int enable);
/* i-th profiled kernel: name (buffer of namelen), total ms, launches; returns
 * BFPSH_EINVAL when i is out of range. */
int bfpsh_profile_get(bfpsh_ctx* ctx, int i, char* name, int namelen, double* total_ms, int* count);
/* Bytes of device memory held by the context. */
int bfpsh_device_bytes(bfpsh_ctx* ctx, size_t* bytes);

/* Replaces the plan creation with `plan_rigor` in the field constructor
 * (bfps/cpp/field.cpp:105-118): chooses the tile shape of each strided pass
 * kind of this context by `rigor` (BFPSH_ESTIMATE / MEASURE / PATIENT, see
 * above).  Measuring allocates one scratch plane for the time it runs and,
 * as FFTW_MEASURE overwrites the arrays it plans for, overwrites the
 * context's work planes (the spectral velocity of compute_velocity); the
 * vorticity and the Runge-Kutta stages are left untouched. */
int bfpsh_set_plan_rigor(bfpsh_ctx* ctx, int rigor);
/* The tile shape in use for pass `kind` (BFPSH_PASS_*): e complex values per
 * thread, b columns per tile, persistent = 1 for the form whose workgroups
 * stay resident and load their next tile while transforming the current
 * one; nshapes = shapes compiled for this box size.  No reference
 * counterpart (FFTW plans are opaque). */
int bfpsh_get_pass_shape(bfpsh_ctx* ctx, int kind, int* e, int* b, int* persistent, int* index, int* nshapes);
/* Forces shape `index` (0 <= index < nshapes; 0 is the built-in one) for
 * pass `kind`; for tests of every shape.  No reference counterpart. */
int bfpsh_set_pass_shape(bfpsh_ctx* ctx, int kind, int index);
/* Replace fftwf_import_wisdom_from_filename / fftwf_export_wisdom_to_filename
 * (bfps/cpp/full_code/main_code.hpp:101-104,145-149) and fftw_forget_wisdom:
 * the process-wide pass wisdom as a text file, one line per (architecture,
 * precision, N, pass kind): "<arch>/<f32|f64>/<N>/<kind> <e> <b> <persistent> <ms> <rigor>".
 * Importing a missing file is not an error (as the reference's start of a
 * new simulation); lines that do not parse are rejected. */
int bfpsh_wisdom_import(const char* path);
int bfpsh_wisdom_export(const char* path);
int bfpsh_wisdom_forget(void);

#ifdef __cplusplus
}
#endif

#endif /* BFPSH_H */
