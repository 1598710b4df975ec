// An NSVEparticles job on the HIP backend written as the reference's C++ is:
// the classes below follow bfps/cpp/full_code/direct_numerical_simulation.cpp,
// NSVE.cpp and NSVEparticles.cpp call for call (initialize, main_loop, step,
// do_stats, write_checkpoint, finalize), and main follows main_code.hpp:121-132.
// The changes: the backend enumerator FFTW -> HIP, the MPI communicator -> an
// int, and the HDF5 C library calls on the run file -> the checkpoint-file
// layer of bfps_hip_h5.hpp (H5Fopen -> h5::File, an hid_t statistics group ->
// hdf5_group).  read_parameters is what the reference's code generator emits
// (parameters/<name> of <simname>.h5).
//
//   nsve_particles_job <simname> [single|double]     (in the run's directory)
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "bfps_hip.hpp"

using namespace bfps_hip;

typedef int MPI_Comm;
static const MPI_Comm MPI_COMM_WORLD = 0;

// ---------------------------------------------------------------- code_base + direct_numerical_simulation
class direct_numerical_simulation {
   public:
    MPI_Comm comm;
    std::string simname;
    int iteration = 0, checkpoint = 0;
    bool stop_code_now = false;
    std::unique_ptr<h5::File> stat_file;
    // parameters of every DNS (DNS.py:139-166)
    int checkpoints_per_file, dealias_type, niter_out, niter_stat, niter_todo, nx, ny, nz;
    double dkx, dky, dkz;

    direct_numerical_simulation(const MPI_Comm COMMUNICATOR, const std::string& simulation_name)
        : comm(COMMUNICATOR), simname(simulation_name) {}
    virtual ~direct_numerical_simulation() {}

    virtual int initialize() = 0;
    virtual int step() = 0;
    virtual int do_stats() = 0;
    virtual int write_checkpoint() = 0;
    virtual int finalize() = 0;

    /* direct_numerical_simulation.cpp:15-20 (hdf5_tools::grow_file_datasets) */
    int grow_file_datasets() {
        const uint64_t rows = uint64_t(niter_todo / niter_stat);
        for (const std::string& kind : stat_file->keys("statistics"))
            for (const std::string& name : stat_file->keys("statistics/" + kind))
                stat_file->grow_rows("statistics/" + kind + "/" + name, rows);
        return 0;
    }
    /* direct_numerical_simulation.cpp:22-55 */
    int read_iteration() {
        h5::File iteration_file(simname + ".h5", 'r');
        iteration = int(iteration_file.read<int64_t>("iteration")[0]);
        checkpoint = int(iteration_file.read<int64_t>("checkpoint")[0]);
        return EXIT_SUCCESS;
    }
    /* direct_numerical_simulation.cpp:57-85 */
    int write_iteration() {
        const int64_t it = iteration, ck = checkpoint;
        stat_file->write<int64_t>("iteration", {}, &it);
        stat_file->write<int64_t>("checkpoint", {}, &ck);
        return EXIT_SUCCESS;
    }
    /* code_base.cpp:15-35 */
    int check_stopping_condition() {
        struct stat file_buffer;
        stop_code_now = stat(("stop_" + simname).c_str(), &file_buffer) == 0;
        return EXIT_SUCCESS;
    }
    /* direct_numerical_simulation.cpp:88-118 */
    int main_loop() {
        const int max_iter = iteration + niter_todo - (iteration % niter_todo);
        for (; iteration < max_iter;) {
            do_stats();
            step();
            if (iteration % niter_out == 0) write_checkpoint();
            check_stopping_condition();
            if (stop_code_now) break;
        }
        do_stats();
        if (iteration % niter_out != 0) write_checkpoint();
        return EXIT_SUCCESS;
    }

   protected:
    // what the generated read_parameters does with hdf5_tools::read_value
    int read_int(h5::File& f, const char* name) { return int(f.read<int64_t>(std::string("parameters/") + name)[0]); }
    double read_double(h5::File& f, const char* name) { return f.read<double>(std::string("parameters/") + name)[0]; }
    virtual int read_parameters() {
        h5::File f(simname + ".h5", 'r');
        checkpoints_per_file = read_int(f, "checkpoints_per_file");
        dealias_type = read_int(f, "dealias_type");
        niter_out = read_int(f, "niter_out");
        niter_stat = read_int(f, "niter_stat");
        niter_todo = read_int(f, "niter_todo");
        nx = read_int(f, "nx");
        ny = read_int(f, "ny");
        nz = read_int(f, "nz");
        dkx = read_double(f, "dkx");
        dky = read_double(f, "dky");
        dkz = read_double(f, "dkz");
        return EXIT_SUCCESS;
    }
};

// ---------------------------------------------------------------- NSVE (NSVE.cpp)
template <typename rnumber>
class NSVE : public direct_numerical_simulation {
   public:
    double dt, famplitude, fk0, fk1;
    int fmode;
    char forcing_type[512];
    int histogram_bins;
    double max_velocity_estimate, max_vorticity_estimate, nu;

    vorticity_equation<rnumber, HIP>* fs = nullptr;
    field<rnumber, HIP, THREE>* tmp_vec_field = nullptr;

    NSVE(const MPI_Comm COMMUNICATOR, const std::string& simulation_name)
        : direct_numerical_simulation(COMMUNICATOR, simulation_name) {}

    int initialize() override {
        this->read_iteration();
        this->read_parameters();
        this->stat_file.reset(new h5::File(this->simname + ".h5", 'a'));
        const int data_file_problem = this->grow_file_datasets();
        if (data_file_problem > 0) {
            std::cerr << data_file_problem << " problems growing file datasets.\ntrying to exit now." << std::endl;
            return EXIT_FAILURE;
        }
        this->fs = new vorticity_equation<rnumber, HIP>(simname.c_str(), nx, ny, nz, dkx, dky, dkz, DEFAULT_HIP_FLAG);
        this->tmp_vec_field = new field<rnumber, HIP, THREE>(nx, ny, nz, this->comm, 0);

        this->fs->checkpoints_per_file = checkpoints_per_file;
        this->fs->nu = nu;
        this->fs->fmode = fmode;
        this->fs->famplitude = famplitude;
        this->fs->fk0 = fk0;
        this->fs->fk1 = fk1;
        strncpy(this->fs->forcing_type, forcing_type, 128);
        this->fs->iteration = this->iteration;
        this->fs->checkpoint = this->checkpoint;

        this->fs->cvorticity->real_space_representation = false;
        this->fs->io_checkpoint();

        if (this->iteration == 0) this->fs->kk->store(*this->stat_file);
        return EXIT_SUCCESS;
    }
    int step() override {
        this->fs->step(this->dt);
        this->iteration = this->fs->iteration;
        return EXIT_SUCCESS;
    }
    int write_checkpoint() override {
        this->fs->io_checkpoint(false);
        this->checkpoint = this->fs->checkpoint;
        this->write_iteration();
        return EXIT_SUCCESS;
    }
    int finalize() override {
        this->stat_file->close();
        delete this->fs;
        delete this->tmp_vec_field;
        return EXIT_SUCCESS;
    }
    int do_stats() override {
        if (!(this->iteration % this->niter_stat == 0)) return EXIT_SUCCESS;
        const hdf5_group stat_group{this->stat_file.get(), "statistics"};

        *tmp_vec_field = fs->cvorticity->get_cdata();
        tmp_vec_field->compute_stats(fs->kk, stat_group, "vorticity", fs->iteration / niter_stat,
                                     max_vorticity_estimate / sqrt(3));

        fs->compute_velocity(fs->cvorticity);
        *tmp_vec_field = fs->cvelocity->get_cdata();
        tmp_vec_field->compute_stats(fs->kk, stat_group, "velocity", fs->iteration / niter_stat,
                                     max_velocity_estimate / sqrt(3));
        return EXIT_SUCCESS;
    }

   protected:
    int read_parameters() override {
        this->direct_numerical_simulation::read_parameters();
        h5::File f(simname + ".h5", 'r');
        dt = read_double(f, "dt");
        famplitude = read_double(f, "famplitude");
        fk0 = read_double(f, "fk0");
        fk1 = read_double(f, "fk1");
        fmode = read_int(f, "fmode");
        std::snprintf(forcing_type, sizeof(forcing_type), "%s", f.read_string("parameters/forcing_type").c_str());
        histogram_bins = read_int(f, "histogram_bins");
        max_velocity_estimate = read_double(f, "max_velocity_estimate");
        max_vorticity_estimate = read_double(f, "max_vorticity_estimate");
        nu = read_double(f, "nu");
        return EXIT_SUCCESS;
    }
};

// ---------------------------------------------------------------- NSVEparticles (NSVEparticles.cpp)
template <typename rnumber>
class NSVEparticles : public NSVE<rnumber> {
   public:
    int niter_part, nparticles, tracers0_integration_steps, tracers0_neighbours, tracers0_smoothness;
    std::unique_ptr<abstract_particles_system<long long int, double>> ps;
    particles_output_hdf5<long long int, double, 3, 3>* particles_output_writer_mpi = nullptr;

    NSVEparticles(const MPI_Comm COMMUNICATOR, const std::string& simulation_name)
        : NSVE<rnumber>(COMMUNICATOR, simulation_name) {}

    int initialize() override {
        this->NSVE<rnumber>::initialize();

        this->ps = particles_system_builder(
            this->fs->cvelocity,                                                    // (field object)
            this->fs->kk,                                                           // (kspace object)
            tracers0_integration_steps,                                             // nb rhs
            (long long int)nparticles,                                              // total particles
            this->fs->get_current_fname(),                                          // particles input filename
            std::string("/tracers0/state/") + std::to_string(this->fs->iteration),  // initial state
            std::string("/tracers0/rhs/") + std::to_string(this->fs->iteration),    // initial rhs
            tracers0_neighbours,                                                    // interpolation neighbours
            tracers0_smoothness,                                                    // smoothness
            this->comm, this->fs->iteration + 1);
        this->particles_output_writer_mpi = new particles_output_hdf5<long long int, double, 3, 3>(
            MPI_COMM_WORLD, "tracers0", nparticles, tracers0_integration_steps);
        return EXIT_SUCCESS;
    }
    int step() override {
        this->fs->compute_velocity(this->fs->cvorticity);
        this->fs->cvelocity->ift();
        this->ps->completeLoop(this->dt);
        this->NSVE<rnumber>::step();
        return EXIT_SUCCESS;
    }
    int write_checkpoint() override {
        this->NSVE<rnumber>::write_checkpoint();
        this->particles_output_writer_mpi->open_file(this->fs->get_current_fname());
        this->particles_output_writer_mpi->save(this->ps->getParticlesPositions(), this->ps->getParticlesRhs(),
                                                this->ps->getParticlesIndexes(), this->ps->getLocalNbParticles(),
                                                this->fs->iteration);
        this->particles_output_writer_mpi->close_file();
        return EXIT_SUCCESS;
    }
    int finalize() override {
        this->NSVE<rnumber>::finalize();
        this->ps.release();
        delete this->particles_output_writer_mpi;
        return EXIT_SUCCESS;
    }
    int do_stats() override {
        this->NSVE<rnumber>::do_stats();
        if (!(this->iteration % this->niter_part == 0)) return EXIT_SUCCESS;

        sample_from_particles_system(*this->tmp_vec_field, this->ps, (this->simname + "_particles.h5"), "tracers0",
                                     "velocity");

        this->fs->compute_Lagrangian_acceleration(this->tmp_vec_field);
        this->tmp_vec_field->ift();
        sample_from_particles_system(*this->tmp_vec_field, this->ps, (this->simname + "_particles.h5"), "tracers0",
                                     "acceleration");
        return EXIT_SUCCESS;
    }

   protected:
    int read_parameters() override {
        this->NSVE<rnumber>::read_parameters();
        h5::File f(this->simname + ".h5", 'r');
        niter_part = this->read_int(f, "niter_part");
        nparticles = this->read_int(f, "nparticles");
        tracers0_integration_steps = this->read_int(f, "tracers0_integration_steps");
        tracers0_neighbours = this->read_int(f, "tracers0_neighbours");
        tracers0_smoothness = this->read_int(f, "tracers0_smoothness");
        return EXIT_SUCCESS;
    }
};

// ---------------------------------------------------------------- main (main_code.hpp:121-132)
template <class DNS>
static int run(const std::string& simname) {
    // pass wisdom around the run, as main_code.hpp:101-104,145-149 does with
    // the FFTW wisdom
    const std::string wisdom = simname + "_hip_wisdom.txt";
    if (bfpsh_wisdom_import(wisdom.c_str()) != BFPSH_OK) std::cerr << "ignoring unreadable " << wisdom << "\n";
    DNS c(MPI_COMM_WORLD, simname);
    if (c.initialize() != EXIT_SUCCESS) return EXIT_FAILURE;
    c.main_loop();
    c.finalize();
    if (bfpsh_wisdom_export(wisdom.c_str()) != BFPSH_OK) std::cerr << "could not write " << wisdom << "\n";
    return EXIT_SUCCESS;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: nsve_particles_job <simname> [single|double]\n";
        return 2;
    }
    const std::string precision = argc > 2 ? argv[2] : "single";
    try {
        return precision == "double" ? run<NSVEparticles<double>>(argv[1]) : run<NSVEparticles<float>>(argv[1]);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
