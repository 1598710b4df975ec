// One (precision, N) instantiation of the launchers; compiled once per pair
// with -DBFPSH_T=float|double -DBFPSH_TN=f32|f64 -DBFPSH_N=<size>.
#include "impl.hpp"

#define BFPSH_CAT3(a, b, c) a##b##_##c
#define BFPSH_NAME(tn, n) BFPSH_CAT3(bfpsh_ops_, tn, n)

const OpsTable* BFPSH_NAME(BFPSH_TN, BFPSH_N)() {
    using I = Impl<BFPSH_T, BFPSH_N>;
    static const OpsTable t = {&I::step, &I::load_projection, &I::velocity_rspace, &I::gradient_rspace,
                               &I::spectral_velocity, &I::cospectrum, &I::field_ift, &I::field_dft,
                               &I::batch_step_desc, &I::pressure_v, &I::lagrangian_acceleration_v, &I::normalize,
                               &I::rspace_stats, &I::compute_vorticity, &I::tune_passes, &I::shapes_of};
    return &t;
}
