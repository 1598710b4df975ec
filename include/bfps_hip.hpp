/* bfps_hip.hpp -- the reference's C++ class surface (field, kspace,
 * vorticity_equation, abstract_particles_system, particles_system_builder)
 * over the C ABI of bfpsh.h, so that a driver written against
 * bfps/cpp/full_code/NSVE.cpp and NSVEparticles.cpp compiles against the HIP
 * backend with the backend enumerator changed from FFTW to HIP.
 *
 * Header-only; link with libbfps_hip.so.  One process drives one GPU; every
 * box (vorticity_equation) owns a device context, and standalone fields
 * built with the reference constructor field(nx, ny, nz, comm, rigor) live
 * on the most recently created box of the same size and precision.
 *
 * Differences from the reference that a caller can observe:
 *   - Data stay on the device.  get_rdata() / get_cdata() download into a
 *     host mirror (reference layouts, real rows padded to nx+2) and return
 *     it; after writing through those pointers call upload() (operator=
 *     from a host array uploads by itself).
 *   - Errors throw bfps_hip::error (the reference mostly ignores them).
 *   - particles_system: compute, move, redistribute, inc_step_idx and
 *     shift_rhs_vectors are fused into completeLoop on the device; calling
 *     them one by one throws.  HDF5 input is replaced by arrays.
 *   - The sequence of NSVEparticles::step (compute_velocity(cvorticity);
 *     cvelocity->ift(); ps->completeLoop(dt); fs->step(dt)) is recognised
 *     and run as one fused step in which the tracers use the velocity of the
 *     first Runge-Kutta stage (the same u(t)); any other call in between
 *     runs the deferred work first.
 */
#ifndef BFPS_HIP_HPP
#define BFPS_HIP_HPP

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfpsh.h"
#include "bfps_hip_h5.hpp"

namespace bfps_hip {

/* bfps/cpp/kspace.hpp:39-40, field_layout.hpp:33 */
enum field_backend { HIP };
enum kspace_dealias_type { TWO_THIRDS, SMOOTH };
enum field_components { ONE, THREE, THREExTHREE };

constexpr int ncomp(field_components fc) { return fc == ONE ? 1 : (fc == THREE ? 3 : 9); }

/* Planner flags of the constructors, as fftw3's FFTW_ESTIMATE /
 * FFTW_MEASURE / FFTW_PATIENT, and the drivers' default of
 * fftw_interface.hpp:30-35 (PATIENT unless USE_HIPESTIMATE is defined):
 * they pick how the strided FFT passes choose their tile shapes
 * (bfpsh_set_plan_rigor). */
constexpr unsigned HIP_ESTIMATE = BFPSH_ESTIMATE, HIP_MEASURE = BFPSH_MEASURE, HIP_PATIENT = BFPSH_PATIENT;
#ifdef USE_HIPESTIMATE
#define DEFAULT_HIP_FLAG bfps_hip::HIP_ESTIMATE
#else
#define DEFAULT_HIP_FLAG bfps_hip::HIP_PATIENT
#endif

class error : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

/* A group of an open HDF5 file (the statistics group of the run file):
 * what the reference's statistics writers take as an hid_t. */
struct hdf5_group {
    h5::File* file;
    std::string path;
    std::string at(const std::string& name) const {
        return (path.empty() || path == "/") ? name : path + "/" + name;
    }
};

namespace detail {

inline void check(bfpsh_ctx* ctx, int rc, const char* what) {
    if (rc != 0) {
        const char* msg = ctx ? bfpsh_last_error(ctx) : "";
        throw error(std::string(what) + " failed: " + (msg ? msg : ""));
    }
}

template <typename rnumber> struct precision;
template <> struct precision<float> { static constexpr int value = BFPSH_FP32; };
template <> struct precision<double> { static constexpr int value = BFPSH_FP64; };

/* One device box plus the deferred work of the NSVEparticles::step sequence. */
struct box {
    bfpsh_ctx* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0, prec = 0;
    bool velocity_ift_pending = false;  // cvelocity->ift() requested
    bool loop_pending = false;          // completeLoop(dt) requested
    double loop_dt = 0.0;

    /* Runs whatever was deferred, in the order it was requested. */
    void flush() {
        if (loop_pending) {
            loop_pending = false;
            velocity_ift_pending = false;
            check(ctx, bfpsh_velocity_rspace(ctx), "velocity_rspace");
            check(ctx, bfpsh_particles_complete_loop(ctx, loop_dt), "completeLoop");
        } else if (velocity_ift_pending) {
            velocity_ift_pending = false;
            check(ctx, bfpsh_velocity_rspace(ctx), "velocity_rspace");
        }
    }
};

inline std::vector<box*>& boxes() {
    static std::vector<box*> registry;
    return registry;
}

inline box* find_box(int nx, int ny, int nz, int prec) {
    auto& r = boxes();
    for (auto it = r.rbegin(); it != r.rend(); ++it)
        if ((*it)->nx == nx && (*it)->ny == ny && (*it)->nz == nz && (*it)->prec == prec) return *it;
    throw error("no vorticity_equation of this size and precision exists to hold the field");
}

enum class role { standalone, vorticity, velocity };

}  // namespace detail

template <field_backend be, kspace_dealias_type dt>
class kspace;

/* field<rnumber, HIP, fc> (bfps/cpp/field.hpp:53-289). */
template <typename rnumber, field_backend be, field_components fc>
class field {
   public:
    typedef rnumber cnumber[2];

    int64_t npoints;
    bool real_space_representation = false;
    int nx, ny, nz;

    /* field(nx, ny, nz, comm, plan_rigor) (field.cpp:37-121): attaches to
     * the box of the same size and precision.  comm and plan rigor are
     * accepted for source compatibility and ignored (one GPU, no plans). */
    template <typename comm_t>
    field(int nx_, int ny_, int nz_, comm_t, unsigned = 0)
        : field(detail::find_box(nx_, ny_, nz_, detail::precision<rnumber>::value), detail::role::standalone) {}

    field(detail::box* b, detail::role r) : npoints(int64_t(b->nx) * b->ny * b->nz), nx(b->nx), ny(b->ny), nz(b->nz),
                                            box_(b), role_(r) {
        if (r == detail::role::standalone)
            detail::check(b->ctx, bfpsh_field_create(b->ctx, ncomp(fc), &id_), "field_create");
    }
    field(const field&) = delete;
    field& operator=(const field&) = delete;
    ~field() {
        if (role_ == detail::role::standalone && id_ >= 0 && box_->ctx) bfpsh_field_destroy(box_->ctx, id_);
    }

    /* field::dft / ift (field.cpp:144-162), unnormalised, in place. */
    void dft() {
        if (role_ != detail::role::standalone) throw error("dft of a solver field: copy it into a standalone field");
        box_->flush();
        detail::check(box_->ctx, bfpsh_field_dft(box_->ctx, id_), "dft");
        real_space_representation = false;
    }
    void ift() {
        if (role_ == detail::role::velocity) {
            // computed from cvorticity when first needed (see compute_velocity)
            box_->velocity_ift_pending = true;
            real_space_representation = true;
            return;
        }
        if (role_ != detail::role::standalone) throw error("ift of the solver vorticity: copy it first");
        box_->flush();
        detail::check(box_->ctx, bfpsh_field_ift(box_->ctx, id_), "ift");
        real_space_representation = true;
    }
    /* field::normalize (field.cpp:891-898): divide by npoints. */
    void normalize() {
        if (role_ == detail::role::standalone) {
            box_->flush();
            detail::check(box_->ctx, bfpsh_field_normalize(box_->ctx, id_), "normalize");
            return;
        }
        const bool r = real_space_representation;
        download();
        for (auto& v : host_) v /= rnumber(npoints);
        real_space_representation = r;
        upload();
    }
    /* field::compute_rspace_stats (field.cpp:661-850) without the HDF5
     * writes: moments [10][nvals] and hist [nbins][nvals] (nvals = 4 for a
     * THREE field, the magnitude last, else ncomp); max_estimate holds nvals
     * values (the reference's max_estimate_vector). */
    void compute_rspace_stats(const std::vector<double>& max_estimate, int nbins, std::vector<double>& moments,
                              std::vector<int64_t>& hist) {
        if (role_ != detail::role::standalone) throw error("rspace stats of a solver field: copy it first");
        if (!real_space_representation) throw error("compute_rspace_stats needs the real-space representation");
        const size_t nvals = ncomp(fc) == 3 ? 4 : ncomp(fc);
        if (max_estimate.size() != nvals) throw error("max_estimate needs one value per statistic");
        box_->flush();
        moments.assign(10 * nvals, 0.0);
        hist.assign(size_t(nbins) * nvals, 0);
        detail::check(box_->ctx, bfpsh_field_rspace_stats(box_->ctx, id_, nbins, max_estimate.data(), moments.data(),
                                                          hist.data()),
                      "rspace_stats");
    }
    /* field::compute_rspace_stats (field.cpp:661-850) into the statistics
     * group: row toffset of moments/<dset_name> [t][10][nvals] and of
     * histograms/<dset_name> [t][nbins][nvals]; nbins is the histogram
     * dataset's second dimension, as there. */
    void compute_rspace_stats(const hdf5_group& group, const std::string& dset_name, uint64_t toffset,
                              const std::vector<double>& max_estimate) {
        const std::vector<uint64_t> hs = group.file->shape(group.at("histograms/" + dset_name));
        if (hs.size() < 2) throw error("histograms/" + dset_name + " has no bin dimension");
        std::vector<double> moments;
        std::vector<int64_t> hist;
        compute_rspace_stats(max_estimate, int(hs[1]), moments, hist);
        group.file->write_rows(group.at("moments/" + dset_name), toffset, 1, moments.data());
        group.file->write_rows(group.at("histograms/" + dset_name), toffset, 1, hist.data());
    }
    /* field::compute_stats (field.cpp:984-1040): real-space statistics and
     * the cospectrum of the field with itself (spectra/<name>_<name>), row
     * toffset of the statistics group; the field ends in the representation
     * it did not start in, with a forward transform normalised, as there. */
    template <kspace_dealias_type dt>
    void compute_stats(kspace<be, dt>* kk, const hdf5_group& group, const std::string& dset_name, uint64_t toffset,
                       double max_estimate) {
        std::vector<double> mev;
        if (fc == ONE) mev.assign(1, max_estimate);
        else if (fc == THREE) mev = {max_estimate, max_estimate, max_estimate, max_estimate * std::sqrt(3.0)};
        else mev.assign(9, max_estimate);
        bool did_rspace = false;
        if (real_space_representation) {
            compute_rspace_stats(group, dset_name, toffset, mev);
            did_rspace = true;
            dft();
            normalize();
        }
        kk->cospectrum(*this, *this, group, dset_name + "_" + dset_name, toffset);
        if (!did_rspace) {
            ift();
            compute_rspace_stats(group, dset_name, toffset, mev);
        }
    }
    /* field::io (field.cpp:164-356): dataset <field_name>/complex/<it> (or
     * /real/<it>, by the current representation) of fname; a complex read
     * is symmetrized, as there.  Real data are stored unpadded, [nz][ny][nx]
     * followed by the component dimensions. */
    void io(const std::string& fname, const std::string& field_name, int iteration, bool read) {
        const std::string dset = field_name + (real_space_representation ? "/real/" : "/complex/") +
                                 std::to_string(iteration);
        const size_t nc = ncomp(fc);
        if (read) {
            h5::File f(fname, 'r');
            const std::vector<rnumber> v = f.read<rnumber>(dset);
            if (real_space_representation) {
                if (v.size() != size_t(npoints) * nc) throw error(dset + " has the wrong size");
                host_.assign(rsize(), rnumber(0));
                for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                    std::memcpy(&host_[zy * (nx + 2) * nc], &v[zy * nx * nc], sizeof(rnumber) * nx * nc);
            } else {
                if (v.size() != csize()) throw error(dset + " has the wrong size");
                host_ = v;
            }
            upload();
            if (!real_space_representation && role_ == detail::role::standalone) symmetrize();
            return;
        }
        download();
        std::vector<uint64_t> shape;
        h5::dtype t{'f', sizeof(rnumber)};
        std::vector<rnumber> dense;
        const rnumber* src = host_.data();
        if (real_space_representation) {
            shape = {uint64_t(nz), uint64_t(ny), uint64_t(nx)};
            dense.resize(size_t(npoints) * nc);
            for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                std::memcpy(&dense[zy * nx * nc], &host_[zy * (nx + 2) * nc], sizeof(rnumber) * nx * nc);
            src = dense.data();
        } else {
            shape = {uint64_t(ny), uint64_t(nz), uint64_t(nx / 2 + 1)};
            t = {'c', 2 * sizeof(rnumber)};
        }
        if (fc == THREE) shape.push_back(3);
        if (fc == THREExTHREE) shape.insert(shape.end(), {3, 3});
        h5::File f(fname, 'a');
        f.write(dset, t, shape, src);
        f.close();
    }
    /* field::symmetrize (field.cpp:900-978). */
    void symmetrize() {
        if (role_ != detail::role::standalone) throw error("symmetrize of a solver field: it is kept symmetric");
        box_->flush();
        detail::check(box_->ctx, bfpsh_field_symmetrize(box_->ctx, id_), "symmetrize");
    }

    /* Host mirror in the reference layouts (field.cpp:85-101), holding the
     * current representation; as in the reference both accessors return
     * the same buffer. */
    rnumber* get_rdata() {
        download();
        return host_.data();
    }
    cnumber* get_cdata() {
        download();
        return reinterpret_cast<cnumber*>(host_.data());
    }
    /* Pushes the host mirror (as last returned by get_rdata/get_cdata) back. */
    void upload() {
        box_->flush();
        const size_t nc = ncomp(fc);
        if (real_space_representation) {
            if (role_ != detail::role::standalone) throw error("real-space upload into a solver field");
            std::vector<rnumber> dense(size_t(npoints) * nc);
            for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                std::memcpy(&dense[zy * nx * nc], &host_[zy * (nx + 2) * nc], sizeof(rnumber) * nx * nc);
            detail::check(box_->ctx, bfpsh_field_put(box_->ctx, id_, 1, dense.data()), "field_put");
        } else if (role_ == detail::role::vorticity) {
            detail::check(box_->ctx, bfpsh_put_cfield(box_->ctx, BFPSH_CVORTICITY, host_.data()), "put_cfield");
        } else if (role_ == detail::role::standalone) {
            detail::check(box_->ctx, bfpsh_field_put(box_->ctx, id_, 0, host_.data()), "field_put");
        } else {
            // the `u` member compute_vorticity reads
            detail::check(box_->ctx, bfpsh_put_cfield(box_->ctx, BFPSH_CVELOCITY, host_.data()), "put_cfield");
        }
    }

    field& operator=(const cnumber* source) {
        real_space_representation = false;
        host_.assign(reinterpret_cast<const rnumber*>(source),
                     reinterpret_cast<const rnumber*>(source) + csize());
        upload();
        return *this;
    }
    field& operator=(const rnumber* source) {
        real_space_representation = true;
        host_.assign(source, source + rsize());
        upload();
        return *this;
    }
    field& operator=(const rnumber value) {
        host_.assign(real_space_representation ? rsize() : csize(), value);
        upload();
        return *this;
    }

    /* index helpers on the host mirror (field.hpp:262-288) */
    ptrdiff_t get_cindex(ptrdiff_t x, ptrdiff_t y, ptrdiff_t z) const { return (y * nz + z) * (nx / 2 + 1) + x; }
    ptrdiff_t get_rindex(ptrdiff_t x, ptrdiff_t y, ptrdiff_t z) const { return (z * ny + y) * (nx + 2) + x; }
    rnumber& rval(ptrdiff_t rindex, unsigned c = 0) { return host_[rindex * ncomp(fc) + c]; }
    rnumber& rval(ptrdiff_t rindex, int c1, int c0) { return host_[(rindex * 3 + c1) * 3 + c0]; }
    rnumber& cval(ptrdiff_t cindex, int imag) { return host_[cindex * 2 + imag]; }
    rnumber& cval(ptrdiff_t cindex, int c, int imag) { return host_[(cindex * ncomp(fc) + c) * 2 + imag]; }
    rnumber& cval(ptrdiff_t cindex, int c1, int c0, int imag) { return host_[((cindex * 3 + c1) * 3 + c0) * 2 + imag]; }

    /* field::RLOOP (field.hpp:234-260) over the host mirror. */
    template <class func_type>
    void RLOOP(func_type expression) {
        for (ptrdiff_t z = 0; z < nz; z++)
            for (ptrdiff_t y = 0; y < ny; y++) {
                ptrdiff_t rindex = (z * ny + y) * (nx + 2);
                for (ptrdiff_t x = 0; x < nx; x++) expression(rindex++, x, y, z);
            }
    }

    int handle() const { return id_; }
    detail::role kind() const { return role_; }
    detail::box* owner() const { return box_; }

   private:
    detail::box* box_;
    detail::role role_;
    int id_ = -1;
    std::vector<rnumber> host_;

    size_t rsize() const { return size_t(nz) * ny * (nx + 2) * ncomp(fc); }
    size_t csize() const { return size_t(ny) * nz * (nx / 2 + 1) * ncomp(fc) * 2; }

    void download() {
        box_->flush();
        const size_t nc = ncomp(fc);
        if (real_space_representation) {
            std::vector<rnumber> dense(size_t(npoints) * nc);
            if (role_ == detail::role::velocity)
                detail::check(box_->ctx, bfpsh_get_rfield(box_->ctx, BFPSH_RVELOCITY, dense.data()), "get_rfield");
            else if (role_ == detail::role::standalone)
                detail::check(box_->ctx, bfpsh_field_get(box_->ctx, id_, 1, dense.data()), "field_get");
            else
                throw error("the solver vorticity is spectral");
            host_.assign(rsize(), rnumber(0));
            for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                std::memcpy(&host_[zy * (nx + 2) * nc], &dense[zy * nx * nc], sizeof(rnumber) * nx * nc);
        } else {
            host_.assign(csize(), rnumber(0));
            if (role_ == detail::role::standalone)
                detail::check(box_->ctx, bfpsh_field_get(box_->ctx, id_, 0, host_.data()), "field_get");
            else
                detail::check(box_->ctx,
                              bfpsh_get_cfield(box_->ctx, role_ == detail::role::vorticity ? BFPSH_CVORTICITY
                                                                                            : BFPSH_CVELOCITY,
                                               host_.data()),
                              "get_cfield");
        }
    }
};

/* kspace<HIP, dt> (bfps/cpp/kspace.hpp:43-195): mode tables of a box. */
template <field_backend be, kspace_dealias_type dt>
class kspace {
   public:
    double dkx, dky, dkz, dk, dk2;
    double kM, kM2;
    std::vector<double> kx, ky, kz;
    std::vector<double> kshell;
    std::vector<int64_t> nshell;
    int nshells;

    kspace(detail::box* b, double DKX, double DKY, double DKZ) : dkx(DKX), dky(DKY), dkz(DKZ), box_(b) {
        double info[4];
        detail::check(b->ctx, bfpsh_kspace_info(b->ctx, info, nullptr, nullptr), "kspace_info");
        kM = info[0];
        kM2 = info[1];
        dk = info[2];
        dk2 = dk * dk;
        nshells = int(info[3]);
        kshell.resize(nshells);
        nshell.resize(nshells);
        detail::check(b->ctx, bfpsh_kspace_info(b->ctx, info, kshell.data(), nshell.data()), "kspace_info");
        for (int i = 0; i <= b->nx / 2; i++) kx.push_back(i * dkx);
        for (int i = 0; i < b->ny; i++) ky.push_back((i <= b->ny / 2 ? i : i - b->ny) * dky);
        for (int i = 0; i < b->nz; i++) kz.push_back((i <= b->nz / 2 ? i : i - b->nz) * dkz);
    }

    /* kspace::cospectrum (kspace.cpp:521-598) of two 3-component spectral
     * fields; out [nshells][3][3].  The reference writes it to HDF5; here
     * the caller receives it. */
    template <typename rnumber>
    void cospectrum(field<rnumber, be, THREE>& a, field<rnumber, be, THREE>& b, double* out) {
        box_->flush();
        using detail::role;
        if (a.kind() == role::standalone && b.kind() == role::standalone) {
            detail::check(box_->ctx, bfpsh_field_cospectrum(box_->ctx, a.handle(), b.handle(), out), "cospectrum");
        } else if (a.kind() != role::standalone && b.kind() != role::standalone) {
            auto sel = [](role r) { return r == role::vorticity ? BFPSH_CVORTICITY : BFPSH_CVELOCITY; };
            detail::check(box_->ctx, bfpsh_cospectrum(box_->ctx, sel(a.kind()), sel(b.kind()), out), "cospectrum");
        } else {
            throw error("cospectrum of a solver field with a standalone field: copy the solver field first");
        }
    }

    /* kspace::cospectrum into the statistics group (kspace.cpp:521-598):
     * row toffset of spectra/<dset_name>, [t][nshells][3][3]. */
    template <typename rnumber>
    void cospectrum(field<rnumber, be, THREE>& a, field<rnumber, be, THREE>& b, const hdf5_group& group,
                    const std::string& dset_name, uint64_t toffset) {
        std::vector<double> spec(size_t(nshells) * 9);
        cospectrum(a, b, spec.data());
        group.file->write_rows(group.at("spectra/" + dset_name), toffset, 1, spec.data());
    }

    /* kspace::store (kspace.cpp:178-230): kspace/kshell, nshell, kM and dk
     * of the run file */
    void store(h5::File& stat_file) {
        stat_file.write<double>("kspace/kshell", {uint64_t(nshells)}, kshell.data());
        stat_file.write<int64_t>("kspace/nshell", {uint64_t(nshells)}, nshell.data());
        stat_file.write<double>("kspace/kM", {}, &kM);
        stat_file.write<double>("kspace/dk", {}, &dk);
    }

    /* CLOOP / CLOOP_K2 / CLOOP_K2_NXMODES (kspace.hpp:117-192) on the host. */
    template <class func_type>
    void CLOOP(func_type expression) {
        ptrdiff_t cindex = 0;
        for (ptrdiff_t y = 0; y < ptrdiff_t(ky.size()); y++)
            for (ptrdiff_t z = 0; z < ptrdiff_t(kz.size()); z++)
                for (ptrdiff_t x = 0; x < ptrdiff_t(kx.size()); x++) expression(cindex++, x, y, z);
    }
    template <class func_type>
    void CLOOP_K2(func_type expression) {
        CLOOP([&](ptrdiff_t c, ptrdiff_t x, ptrdiff_t y, ptrdiff_t z) {
            expression(c, x, y, z, kx[x] * kx[x] + ky[y] * ky[y] + kz[z] * kz[z]);
        });
    }
    template <class func_type>
    void CLOOP_K2_NXMODES(func_type expression) {
        CLOOP([&](ptrdiff_t c, ptrdiff_t x, ptrdiff_t y, ptrdiff_t z) {
            expression(c, x, y, z, kx[x] * kx[x] + ky[y] * ky[y] + kz[z] * kz[z], x == 0 ? 1 : 2);
        });
    }

   private:
    detail::box* box_;
};

/* vorticity_equation<rnumber, HIP> (bfps/cpp/vorticity_equation.hpp:46-134). */
template <typename rnumber, field_backend be>
class vorticity_equation {
   public:
    std::string name;
    int iteration = 0, checkpoint = 0, checkpoints_per_file = 1;
    /* physics members written by NSVE::initialize (NSVE.cpp:46-52) */
    double nu = 0.1;
    int fmode = 1;
    double famplitude = 1.0, fk0 = 2.0, fk1 = 4.0;
    char forcing_type[128];

    field<rnumber, be, THREE>* cvorticity;
    field<rnumber, be, THREE>* cvelocity;
    kspace<be, SMOOTH>* kk;

    /* (vorticity_equation.cpp:110-156); dealias and device extend it. */
    vorticity_equation(const char* NAME, int nx, int ny, int nz, double DKX = 1.0, double DKY = 1.0,
                       double DKZ = 1.0, unsigned FFTW_PLAN_RIGOR = HIP_MEASURE, int dealias = BFPSH_SMOOTH,
                       int device = 0)
        : name(NAME) {
        std::strncpy(forcing_type, "linear", sizeof(forcing_type));
        box_.nx = nx;
        box_.ny = ny;
        box_.nz = nz;
        box_.prec = detail::precision<rnumber>::value;
        detail::check(nullptr,
                      bfpsh_create(&box_.ctx, nx, ny, nz, DKX, DKY, DKZ, box_.prec, dealias, device),
                      "vorticity_equation (unsupported size or no GPU)");
        detail::check(box_.ctx, bfpsh_set_plan_rigor(box_.ctx, int(FFTW_PLAN_RIGOR)), "vorticity_equation plan rigor");
        detail::boxes().push_back(&box_);
        cvorticity = new field<rnumber, be, THREE>(&box_, detail::role::vorticity);
        cvelocity = new field<rnumber, be, THREE>(&box_, detail::role::velocity);
        kk = new kspace<be, SMOOTH>(&box_, DKX, DKY, DKZ);
    }
    vorticity_equation(const vorticity_equation&) = delete;
    vorticity_equation& operator=(const vorticity_equation&) = delete;
    ~vorticity_equation() {
        delete kk;
        delete cvelocity;
        delete cvorticity;
        auto& r = detail::boxes();
        for (auto it = r.begin(); it != r.end(); ++it)
            if (*it == &box_) {
                r.erase(it);
                break;
            }
        bfpsh_destroy(box_.ctx);
        box_.ctx = nullptr;
    }

    /* vorticity_equation::step (vorticity_equation.cpp:358-439). */
    void step(double dt) {
        push_physics();
        sync_iteration();
        if (box_.loop_pending && box_.loop_dt == dt) {
            box_.loop_pending = false;
            box_.velocity_ift_pending = false;
            detail::check(box_.ctx, bfpsh_nsve_particles_step(box_.ctx, dt), "NSVEparticles step");
        } else {
            box_.flush();
            detail::check(box_.ctx, bfpsh_step(box_.ctx, dt), "step");
        }
        detail::check(box_.ctx, bfpsh_get_iteration(box_.ctx, &iteration), "get_iteration");
        cvelocity->real_space_representation = false;
    }

    /* compute_velocity(cvorticity) (vorticity_equation.cpp:206-239): the
     * velocity is produced from cvorticity on demand (cvelocity->ift(),
     * get_cdata(), sampling), so this only records the request. */
    void compute_velocity(field<rnumber, be, THREE>* vorticity) {
        if (vorticity != cvorticity) throw error("compute_velocity is supported for cvorticity");
        cvelocity->real_space_representation = false;
    }

    /* get_current_fname (vorticity_equation.hpp:98-104) */
    std::string get_current_fname() const { return name + "_checkpoint_" + std::to_string(checkpoint) + ".h5"; }

    /* update_checkpoint (vorticity_equation.cpp:49-108): the next file once
     * the current one holds checkpoints_per_file fields and not this
     * iteration's; a missing file is created with the group
     * vorticity/complex. */
    void update_checkpoint() {
        const std::string fname = get_current_fname();
        std::FILE* probe = std::fopen(fname.c_str(), "rb");
        if (probe) {
            std::fclose(probe);
            h5::File f(fname, 'r');
            const std::vector<std::string> stored = f.exists("vorticity/complex") ? f.keys("vorticity/complex")
                                                                                    : std::vector<std::string>();
            bool here = false;
            for (const std::string& k : stored) here = here || k == std::to_string(iteration);
            if (int(stored.size()) >= checkpoints_per_file && !here) checkpoint++;
        } else {
            h5::File f(fname, 'w');
            f.create_group("vorticity/complex");
            f.close();
        }
    }

    /* io_checkpoint (vorticity_equation.hpp:107-128): read = true loads
     * vorticity/complex/<iteration> of the current checkpoint file, then
     * symmetrize, low_pass(kM), force_divfree; read = false updates the
     * checkpoint index and writes cvorticity there. */
    void io_checkpoint(bool read = true) {
        if (cvorticity->real_space_representation) throw error("io_checkpoint of a real-space vorticity");
        if (!read) update_checkpoint();
        const std::string fname = get_current_fname();
        if (read) {
            h5::File f(fname, 'r');
            const std::vector<rnumber> v = f.read<rnumber>("vorticity/complex/" + std::to_string(iteration));
            if (v.size() != size_t(box_.ny) * box_.nz * (box_.nx / 2 + 1) * 3 * 2)
                throw error("vorticity/complex/" + std::to_string(iteration) + " has the wrong size");
            io_checkpoint(true, reinterpret_cast<const typename field<rnumber, be, THREE>::cnumber*>(v.data()));
        } else {
            cvorticity->io(fname, "vorticity", iteration, false);
        }
    }
    /* the read of io_checkpoint from an array in the complex layout
     * instead of the file (no reference counterpart) */
    void io_checkpoint(bool read, const typename field<rnumber, be, THREE>::cnumber* data) {
        if (!read) throw error("io_checkpoint(false, data): call io_checkpoint(false)");
        if (!data) throw error("io_checkpoint(true, data) needs data");
        *cvorticity = data;
        detail::check(box_.ctx, bfpsh_load_projection(box_.ctx), "io_checkpoint");
    }

    /* compute_vorticity (vorticity_equation.cpp:171-204): cvorticity = i k x
     * cvelocity (k^2 <= kM^2), symmetrized; cvelocity is what was last
     * assigned to it (*cvelocity = data) */
    void compute_vorticity() {
        box_.flush();
        detail::check(box_.ctx, bfpsh_compute_vorticity(box_.ctx), "compute_vorticity");
        cvorticity->real_space_representation = false;
    }
    /* impose_zero_modes (vorticity_equation.cpp:39-46) */
    void impose_zero_modes() {
        box_.flush();
        detail::check(box_.ctx, bfpsh_impose_zero_modes(box_.ctx), "impose_zero_modes");
    }

    /* compute_pressure (vorticity_equation.cpp:441-520) of the velocity of
     * cvorticity into a standalone ONE field (spectral). */
    void compute_pressure(field<rnumber, be, ONE>* pressure) {
        push_physics();
        box_.flush();
        if (pressure->handle() < 0) throw error("compute_pressure needs a standalone field");
        detail::check(box_.ctx, bfpsh_compute_pressure(box_.ctx, pressure->handle()), "compute_pressure");
        pressure->real_space_representation = false;
    }
    /* compute_Lagrangian_acceleration (vorticity_equation.cpp:522-578) into a
     * standalone THREE field (spectral). */
    void compute_Lagrangian_acceleration(field<rnumber, be, THREE>* acceleration) {
        push_physics();
        box_.flush();
        if (acceleration->handle() < 0) throw error("compute_Lagrangian_acceleration needs a standalone field");
        detail::check(box_.ctx, bfpsh_compute_lagrangian_acceleration(box_.ctx, acceleration->handle()),
                      "compute_lagrangian_acceleration");
        acceleration->real_space_representation = false;
    }

    /* compute_gradient(kk, cvelocity, grad) + grad->ift() into the
     * device's real-space gradient (field.cpp:1042-1101). */
    void velocity_gradient_rspace() {
        box_.flush();
        detail::check(box_.ctx, bfpsh_gradient_rspace(box_.ctx), "gradient_rspace");
    }

    void synchronize() {
        box_.flush();
        detail::check(box_.ctx, bfpsh_synchronize(box_.ctx), "synchronize");
    }

    detail::box* get_box() { return &box_; }

   private:
    detail::box box_;

    void push_physics() {
        int f;
        if (!std::strcmp(forcing_type, "linear")) f = BFPSH_FORCING_LINEAR;
        else if (!std::strcmp(forcing_type, "Kolmogorov")) f = BFPSH_FORCING_KOLMOGOROV;
        else if (!std::strcmp(forcing_type, "fixed_energy_injection_rate")) f = BFPSH_FORCING_FIXED_ENERGY;
        else if (!std::strcmp(forcing_type, "none") || !forcing_type[0]) f = BFPSH_FORCING_NONE;
        else throw error(std::string("unknown forcing_type ") + forcing_type);
        detail::check(box_.ctx, bfpsh_set_physics(box_.ctx, nu, f, famplitude, fk0, fk1, fmode), "set_physics");
    }
    void sync_iteration() { detail::check(box_.ctx, bfpsh_set_iteration(box_.ctx, iteration), "set_iteration"); }
};

/* abstract_particles_system<partsize_t, real_number>
 * (bfps/cpp/particles/abstract_particles_system.hpp:12-55). */
template <class partsize_t, class real_number>
class abstract_particles_system {
   public:
    virtual ~abstract_particles_system() {}
    virtual void compute() = 0;
    virtual void move(const real_number dt) = 0;
    virtual void redistribute() = 0;
    virtual void inc_step_idx() = 0;
    virtual void shift_rhs_vectors() = 0;
    virtual void completeLoop(const real_number dt) = 0;
    virtual const real_number* getParticlesPositions() const = 0;
    virtual const std::unique_ptr<real_number[]>* getParticlesRhs() const = 0;
    virtual const partsize_t* getParticlesIndexes() const = 0;
    virtual partsize_t getLocalNbParticles() const = 0;
    virtual partsize_t getGlobalNbParticles() const = 0;
    virtual int getNbRhs() const = 0;
    virtual int get_step_idx() const = 0;
    virtual void sample_compute_field(const field<float, HIP, ONE>&, real_number[]) = 0;
    virtual void sample_compute_field(const field<float, HIP, THREE>&, real_number[]) = 0;
    virtual void sample_compute_field(const field<float, HIP, THREExTHREE>&, real_number[]) = 0;
    virtual void sample_compute_field(const field<double, HIP, ONE>&, real_number[]) = 0;
    virtual void sample_compute_field(const field<double, HIP, THREE>&, real_number[]) = 0;
    virtual void sample_compute_field(const field<double, HIP, THREExTHREE>&, real_number[]) = 0;
};

/* The device tracer set of one box (particles_system.hpp:22-217 on one rank). */
template <class partsize_t>
class particles_system final : public abstract_particles_system<partsize_t, double> {
   public:
    particles_system(detail::box* b, int nsteps, partsize_t np, int n, int m, int step_idx)
        : box_(b), nsteps_(nsteps), np_(np) {
        detail::check(b->ctx, bfpsh_particles_create(b->ctx, int64_t(np), n, m, nsteps, step_idx),
                      "particles_system_builder");
    }

    void compute() override { fused("compute"); }
    void move(const double) override { fused("move"); }
    void redistribute() override { fused("redistribute"); }
    void inc_step_idx() override { fused("inc_step_idx"); }
    void shift_rhs_vectors() override { fused("shift_rhs_vectors"); }

    /* particles_system::completeLoop (particles_system.hpp:210-217). */
    void completeLoop(const double dt) override {
        if (box_->velocity_ift_pending && !box_->loop_pending) {
            box_->loop_pending = true;  // fused into the next step(dt) when dt matches
            box_->loop_dt = dt;
            return;
        }
        box_->flush();
        detail::check(box_->ctx, bfpsh_particles_complete_loop(box_->ctx, dt), "completeLoop");
    }

    const double* getParticlesPositions() const override {
        download();
        return pos_.data();
    }
    const std::unique_ptr<double[]>* getParticlesRhs() const override {
        download();
        return rhs_.data();
    }
    const partsize_t* getParticlesIndexes() const override {
        download();
        return idx_.data();
    }
    partsize_t getLocalNbParticles() const override { return np_; }
    partsize_t getGlobalNbParticles() const override { return np_; }
    int getNbRhs() const override { return nsteps_; }
    int get_step_idx() const override {
        box_->flush();
        int s = 0;
        detail::check(box_->ctx, bfpsh_particles_get(box_->ctx, nullptr, nullptr, nullptr, &s), "get_step_idx");
        return s;
    }

    /* State as particles_input_hdf5 reads it: positions [np][3], rhs
     * [nsteps][np][3] (may be null), global indexes (may be null). */
    void put(const double* pos, const double* rhs = nullptr, const int64_t* idx = nullptr) {
        box_->flush();
        detail::check(box_->ctx, bfpsh_particles_put(box_->ctx, pos, rhs, idx), "particles put");
    }

    void sample_compute_field(const field<float, HIP, ONE>& f, double out[]) override { sample(f, out); }
    void sample_compute_field(const field<float, HIP, THREE>& f, double out[]) override { sample(f, out); }
    void sample_compute_field(const field<float, HIP, THREExTHREE>& f, double out[]) override { sample(f, out); }
    void sample_compute_field(const field<double, HIP, ONE>& f, double out[]) override { sample(f, out); }
    void sample_compute_field(const field<double, HIP, THREE>& f, double out[]) override { sample(f, out); }
    void sample_compute_field(const field<double, HIP, THREExTHREE>& f, double out[]) override { sample(f, out); }

    /* Velocity gradient of the box (after vorticity_equation::
     * velocity_gradient_rspace) at the tracers, [np][9]. */
    void sample_velocity_gradient(double out[]) {
        box_->flush();
        detail::check(box_->ctx, bfpsh_particles_sample(box_->ctx, BFPSH_RGRADIENT, out), "sample gradient");
    }

   private:
    detail::box* box_;
    int nsteps_;
    partsize_t np_;
    mutable std::vector<double> pos_;
    mutable std::vector<std::unique_ptr<double[]>> rhs_;
    mutable std::vector<partsize_t> idx_;

    [[noreturn]] static void fused(const char* what) {
        throw error(std::string(what) + " is fused into completeLoop on the device");
    }

    void download() const {
        box_->flush();
        std::vector<double> rhs(size_t(nsteps_) * np_ * 3);
        std::vector<int64_t> idx(np_);
        pos_.assign(size_t(np_) * 3, 0.0);
        detail::check(box_->ctx, bfpsh_particles_get(box_->ctx, pos_.data(), rhs.data(), idx.data(), nullptr),
                      "particles get");
        rhs_.clear();
        for (int s = 0; s < nsteps_; s++) {
            rhs_.emplace_back(new double[size_t(np_) * 3]);
            std::memcpy(rhs_.back().get(), &rhs[size_t(s) * np_ * 3], sizeof(double) * np_ * 3);
        }
        idx_.assign(idx.begin(), idx.end());
    }

    template <typename rnumber, field_components fc>
    void sample(const field<rnumber, HIP, fc>& f, double out[]) {
        if (f.owner() != box_) throw error("sampled field lives on another box");
        box_->flush();
        if (f.kind() == detail::role::velocity) {
            if (!f.real_space_representation) throw error("sample of cvelocity before cvelocity->ift()");
            detail::check(box_->ctx, bfpsh_particles_sample(box_->ctx, BFPSH_RVELOCITY, out), "sample velocity");
        } else if (f.kind() == detail::role::standalone) {
            detail::check(box_->ctx, bfpsh_particles_sample_field(box_->ctx, f.handle(), out), "sample");
        } else {
            throw error("the solver vorticity is spectral");
        }
    }
};

/* particles_system_builder (bfps/cpp/particles/particles_system_builder.hpp:238-257)
 * with the HDF5 datasets replaced by arrays: state [np][3], rhs
 * [nsteps][np][3] (may be null). */
template <typename rnumber, typename comm_t>
std::unique_ptr<abstract_particles_system<long long int, double>> particles_system_builder(
    field<rnumber, HIP, THREE>* fs_field, kspace<HIP, SMOOTH>*, int nsteps, long long int nparticles,
    const double* state, const double* rhs, int interpolation_size, int spline_mode, comm_t, int in_current_iteration) {
    auto* ps = new particles_system<long long int>(fs_field->owner(), nsteps, nparticles, interpolation_size,
                                                   spline_mode, in_current_iteration);
    std::unique_ptr<abstract_particles_system<long long int, double>> out(ps);
    if (state) ps->put(state, rhs, nullptr);
    return out;
}

/* particles_system_builder (particles_system_builder.hpp:238-257): the
 * tracer state [np][3] and history [nsteps][np][3] are read from datasets
 * state_dset and rhs_dset of fname (particles_input_hdf5; tracer i of the
 * file gets global index i).  A tracer count or history length that
 * differs from the parameters throws here (the reference constructs that
 * error without throwing it, particles_system_builder.hpp:217-225). */
template <typename rnumber, typename comm_t>
std::unique_ptr<abstract_particles_system<long long int, double>> particles_system_builder(
    field<rnumber, HIP, THREE>* fs_field, kspace<HIP, SMOOTH>* fs_kk, int nsteps, long long int nparticles,
    const std::string& fname, const std::string& state_dset, const std::string& rhs_dset, int interpolation_size,
    int spline_mode, comm_t comm, int in_current_iteration) {
    h5::File f(fname, 'r');
    const std::vector<uint64_t> ss = f.shape(state_dset), rs = f.shape(rhs_dset);
    if (ss.empty() || ss.back() != 3) throw error(state_dset + " is not [..][3]");
    uint64_t np_file = 1;
    for (size_t i = 0; i + 1 < ss.size(); i++) np_file *= ss[i];
    if ((long long int)np_file != nparticles)
        throw error("Nb particles is " + std::to_string(nparticles) + " in the parameters but " +
                    std::to_string(np_file) + " in the particles file.");
    if (rs.empty() || (long long int)rs[0] != nsteps)
        throw error("Nb steps is " + std::to_string(nsteps) + " in the parameters but " +
                    std::to_string(rs.empty() ? 0 : rs[0]) + " in the particles file.");
    const std::vector<double> state = f.read<double>(state_dset);
    const std::vector<double> rhs = f.read<double>(rhs_dset);
    if (rhs.size() != size_t(nsteps) * np_file * 3) throw error(rhs_dset + " has the wrong size");
    return particles_system_builder(fs_field, fs_kk, nsteps, nparticles, state.data(), rhs.data(), interpolation_size,
                                    spline_mode, comm, in_current_iteration);
}

/* particles_output_hdf5<partsize_t, real_number, 3, 3>
 * (particles/particles_output_hdf5.hpp, abstract_particles_output.hpp:144-190):
 * open_file creates <species>/state and <species>/rhs when missing; save
 * writes <species>/state/<it> [np][3] and <species>/rhs/<it>
 * [nb_rhs][np][3], tracers in global-index order. */
template <class partsize_t, class real_number, int size_particle_positions, int size_particle_rhs>
class particles_output_hdf5 {
   public:
    template <typename comm_t>
    particles_output_hdf5(comm_t, const std::string& ps_name, partsize_t total_nb_particles, int nb_rhs,
                          bool = false)
        : name_(ps_name), np_(total_nb_particles), nb_rhs_(nb_rhs) {}

    int open_file(const std::string& filename) {
        file_.reset(new h5::File(filename, 'a'));
        file_->create_group(name_ + "/state");
        file_->create_group(name_ + "/rhs");
        return 0;
    }
    void save(const real_number positions[], const std::unique_ptr<real_number[]> rhs[], const partsize_t index[],
              partsize_t nb_particles, int idx_time_step) {
        if (!file_) throw error("particles_output_hdf5::save before open_file");
        if (nb_particles != np_) throw error("particles_output_hdf5: one rank holds every tracer here");
        std::vector<partsize_t> order(static_cast<size_t>(nb_particles));
        for (partsize_t i = 0; i < nb_particles; i++) order[size_t(i)] = i;
        std::sort(order.begin(), order.end(), [&](partsize_t a, partsize_t b) { return index[a] < index[b]; });
        std::vector<real_number> st(size_t(nb_particles) * size_particle_positions);
        std::vector<real_number> rh(size_t(nb_rhs_) * nb_particles * size_particle_rhs);
        for (partsize_t i = 0; i < nb_particles; i++) {
            const partsize_t src = order[size_t(i)];
            for (int c = 0; c < size_particle_positions; c++)
                st[size_t(i) * size_particle_positions + c] = positions[src * size_particle_positions + c];
            for (int r = 0; r < nb_rhs_; r++)
                for (int c = 0; c < size_particle_rhs; c++)
                    rh[(size_t(r) * nb_particles + i) * size_particle_rhs + c] = rhs[r][src * size_particle_rhs + c];
        }
        const std::string it = std::to_string(idx_time_step);
        file_->write<real_number>(name_ + "/state/" + it, {uint64_t(nb_particles), uint64_t(size_particle_positions)},
                                  st.data());
        file_->write<real_number>(name_ + "/rhs/" + it,
                                  {uint64_t(nb_rhs_), uint64_t(nb_particles), uint64_t(size_particle_rhs)}, rh.data());
    }
    int close_file() {
        if (file_) file_->close();
        file_.reset();
        return 0;
    }

   private:
    std::string name_;
    partsize_t np_;
    int nb_rhs_;
    std::unique_ptr<h5::File> file_;
};

/* sample_from_particles_system (particles/particles_sampling.hpp:15-50): the
 * field at the tracers into <parent>/<fname>/<step_idx> of filename,
 * [1][np][ncomp] in global-index order; an existing dataset is kept. */
template <class partsize_t, class particles_rnumber, class rnumber, field_components fc>
void sample_from_particles_system(const field<rnumber, HIP, fc>& in_field,
                                  std::unique_ptr<abstract_particles_system<partsize_t, particles_rnumber>>& ps,
                                  const std::string& filename, const std::string& parent_groupname,
                                  const std::string& fname) {
    const std::string dset = parent_groupname + "/" + fname + "/" + std::to_string(ps->get_step_idx());
    {
        std::FILE* probe = std::fopen(filename.c_str(), "rb");
        if (probe) {
            std::fclose(probe);
            h5::File f(filename, 'r');
            if (f.exists(dset)) return;
        }
    }
    const partsize_t np = ps->getLocalNbParticles();
    const int nc = ncomp(fc);
    std::vector<particles_rnumber> sample(size_t(np) * nc, 0);
    ps->sample_compute_field(in_field, sample.data());  // rows in global-index order
    h5::File f(filename, 'a');
    f.write<particles_rnumber>(dset, {1, uint64_t(np), uint64_t(nc)}, sample.data());
    f.close();
}

/* compute_gradient (field.cpp:1042-1101) of standalone fields of one box:
 * ONE -> THREE or THREE -> THREExTHREE, dst[j ncomp(src) + i] = i k_j
 * src_i for k^2 < kM^2 (dst spectral). */
template <typename rnumber, field_backend be, field_components fc1, field_components fc2, kspace_dealias_type dt>
int compute_gradient(kspace<be, dt>*, field<rnumber, be, fc1>* src, field<rnumber, be, fc2>* dst) {
    static_assert((fc1 == ONE && fc2 == THREE) || (fc1 == THREE && fc2 == THREExTHREE), "gradient shapes");
    if (src->owner() != dst->owner() || src->handle() < 0 || dst->handle() < 0)
        throw error("compute_gradient needs two standalone fields of one box");
    if (src->real_space_representation) throw error("compute_gradient of a real-space field");
    src->owner()->flush();
    detail::check(src->owner()->ctx, bfpsh_field_compute_gradient(src->owner()->ctx, src->handle(), dst->handle()),
                  "compute_gradient");
    dst->real_space_representation = false;
    return 0;
}

}  // namespace bfps_hip

#endif
