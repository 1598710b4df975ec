// A driver written the way bfps/cpp/full_code/NSVE.cpp and NSVEparticles.cpp
// drive the solver (initialize, then step = compute_velocity; cvelocity->ift;
// completeLoop; fs->step), on the HIP backend's C++ class surface.
//
//   nsve_particles <in.bin> <out.bin>
// in.bin : int32 n, nsteps_tracer, neighbours, smoothness, niter, iteration;
//          int64 np; double nu, famplitude, fk0, fk1, dt;
//          float vorticity {re,im} [n][n][n/2+1][3]; double state [np][3];
//          double rhs [nsteps][np][3]
// out.bin: float vorticity after niter steps; double state; double rhs;
//          int32 iteration, step_idx; double velocity samples [np][3];
//          double cospectrum [nshells][3][3] (velocity, velocity)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bfps_hip.hpp"

using namespace bfps_hip;

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
    if (fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}
template <typename T>
static void wr(FILE* f, const T* p, size_t n) {
    fwrite(p, sizeof(T), n, f);
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int hdr[6];
    rd(in, hdr, 6);
    const int n = hdr[0], nsteps = hdr[1], nb = hdr[2], sm = hdr[3], niter = hdr[4], it0 = hdr[5];
    long long np;
    rd(in, &np, 1);
    double par[5];
    rd(in, par, 5);
    const size_t csz = size_t(n) * n * (n / 2 + 1) * 3;
    std::vector<float> w(csz * 2);
    std::vector<double> state(np * 3), rhs(size_t(nsteps) * np * 3);
    rd(in, w.data(), w.size());
    rd(in, state.data(), state.size());
    rd(in, rhs.data(), rhs.size());
    fclose(in);

    try {
        // NSVE::initialize (NSVE.cpp:35-60)
        auto* fs = new vorticity_equation<float, HIP>("nsve", n, n, n, 1.0, 1.0, 1.0, 0);
        auto* tmp_vec_field = new field<float, HIP, THREE>(n, n, n, 0, 0);
        fs->nu = par[0];
        fs->fmode = 1;
        fs->famplitude = par[1];
        fs->fk0 = par[2];
        fs->fk1 = par[3];
        std::snprintf(fs->forcing_type, 128, "%s", "linear");
        fs->iteration = it0;
        fs->cvorticity->real_space_representation = false;
        fs->io_checkpoint(true, reinterpret_cast<const float(*)[2]>(w.data()));
        // NSVEparticles::initialize (NSVEparticles.cpp:10-22)
        auto ps = particles_system_builder(fs->cvelocity, fs->kk, nsteps, np, state.data(), rhs.data(), nb, sm, 0,
                                           fs->iteration + 1);
        const double dt = par[4];
        for (int i = 0; i < niter; i++) {
            // NSVEparticles::step (NSVEparticles.cpp:33-41)
            fs->compute_velocity(fs->cvorticity);
            fs->cvelocity->ift();
            ps->completeLoop(dt);
            fs->step(dt);
        }
        // NSVE::do_stats-like sampling: copy of the velocity, its ift, samples
        *tmp_vec_field = fs->cvelocity->get_cdata();
        std::vector<double> spec(size_t(fs->kk->nshells) * 9);
        fs->kk->cospectrum(*tmp_vec_field, *tmp_vec_field, spec.data());
        tmp_vec_field->ift();
        std::vector<double> samples(np * 3);
        ps->sample_compute_field(*tmp_vec_field, samples.data());

        FILE* out = fopen(argv[2], "wb");
        if (!out) return 2;
        const float(*cw)[2] = fs->cvorticity->get_cdata();
        wr(out, &cw[0][0], w.size());
        wr(out, ps->getParticlesPositions(), np * 3);
        for (int s = 0; s < nsteps; s++) wr(out, ps->getParticlesRhs()[s].get(), np * 3);
        int tail[2] = {fs->iteration, ps->get_step_idx()};
        wr(out, tail, 2);
        wr(out, samples.data(), samples.size());
        wr(out, spec.data(), spec.size());
        fclose(out);
        ps.reset();
        delete tmp_vec_field;
        delete fs;
    } catch (const error& e) {
        fprintf(stderr, "bfps_hip error: %s\n", e.what());
        return 1;
    }
    return 0;
}
