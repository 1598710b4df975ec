"""ctypes binding of the C ABI declared in include/bfpsh.h.

The library is ``bfps_amd/libbfps_hip.so`` (built by ``bfps_amd/build.py``).
There is no fallback: if the library is missing or no HIP device is usable,
loading or creating a context raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("BFPSH_LIB") or os.path.join(HERE, "libbfps_hip.so")

FP32, FP64 = 0, 1
TWO_THIRDS, SMOOTH = 0, 1
FORCING = {"none": 0, "linear": 1, "Kolmogorov": 2, "fixed_energy_injection_rate": 3}
CVORTICITY, CVELOCITY, RVELOCITY, RGRADIENT = 0, 1, 5, 6
# plan rigor (the reference's FFTW planner flags) and strided pass kinds
ESTIMATE, MEASURE, PATIENT = 0, 1, 2
PLAN_RIGOR = {"FFTW_ESTIMATE": ESTIMATE, "FFTW_MEASURE": MEASURE, "FFTW_PATIENT": PATIENT,
              ESTIMATE: ESTIMATE, MEASURE: MEASURE, PATIENT: PATIENT}
PASS_KINDS = {"z_inv": 0, "y_inv": 1, "z_fwd": 2, "y_fwd": 3, "bs": 4, "y3": 5}

# every exported entry point: name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_LP = ctypes.POINTER(ctypes.c_int64)
_IP = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "bfpsh_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _D, _D, _D, _I, _I, _I]),
    "bfpsh_destroy": (_I, [_P]),
    "bfpsh_last_error": (ctypes.c_char_p, [_P]),
    "bfpsh_set_physics": (_I, [_P, _D, _I, _D, _D, _D, _I]),
    "bfpsh_kspace_info": (_I, [_P, _DP, _DP, _LP]),
    "bfpsh_put_cfield": (_I, [_P, _I, _P]),
    "bfpsh_get_cfield": (_I, [_P, _I, _P]),
    "bfpsh_get_rfield": (_I, [_P, _I, _P]),
    "bfpsh_load_projection": (_I, [_P]),
    "bfpsh_compute_vorticity": (_I, [_P]),
    "bfpsh_impose_zero_modes": (_I, [_P]),
    "bfpsh_field_compute_gradient": (_I, [_P, _I, _I]),
    "bfpsh_step": (_I, [_P, _D]),
    "bfpsh_velocity_rspace": (_I, [_P]),
    "bfpsh_gradient_rspace": (_I, [_P]),
    "bfpsh_cospectrum": (_I, [_P, _I, _I, _DP]),
    "bfpsh_field_create": (_I, [_P, _I, _IP]),
    "bfpsh_field_destroy": (_I, [_P, _I]),
    "bfpsh_field_ift": (_I, [_P, _I]),
    "bfpsh_field_dft": (_I, [_P, _I]),
    "bfpsh_field_symmetrize": (_I, [_P, _I]),
    "bfpsh_field_put": (_I, [_P, _I, _I, _P]),
    "bfpsh_field_get": (_I, [_P, _I, _I, _P]),
    "bfpsh_field_copy_from": (_I, [_P, _I, _I]),
    "bfpsh_field_cospectrum": (_I, [_P, _I, _I, _DP]),
    "bfpsh_field_cospectrum_noalloc": (_I, [_P, _I, _I, _DP]),
    "bfpsh_set_plan_rigor": (_I, [_P, _I]),
    "bfpsh_get_pass_shape": (_I, [_P, _I, _IP, _IP, _IP, _IP, _IP]),
    "bfpsh_set_pass_shape": (_I, [_P, _I, _I]),
    "bfpsh_wisdom_import": (_I, [ctypes.c_char_p]),
    "bfpsh_wisdom_export": (_I, [ctypes.c_char_p]),
    "bfpsh_wisdom_forget": (_I, []),
    "bfpsh_field_normalize": (_I, [_P, _I]),
    "bfpsh_field_rspace_stats": (_I, [_P, _I, _I, _DP, _DP, _LP]),
    "bfpsh_compute_pressure": (_I, [_P, _I]),
    "bfpsh_compute_lagrangian_acceleration": (_I, [_P, _I]),
    "bfpsh_particles_create": (_I, [_P, _L, _I, _I, _I, _I]),
    "bfpsh_particles_put": (_I, [_P, _DP, _DP, _LP]),
    "bfpsh_particles_get": (_I, [_P, _DP, _DP, _LP, _IP]),
    "bfpsh_particles_complete_loop": (_I, [_P, _D]),
    "bfpsh_particles_sample": (_I, [_P, _I, _DP]),
    "bfpsh_particles_sample_field": (_I, [_P, _I, _DP]),
    "bfpsh_nsve_particles_step": (_I, [_P, _D]),
    "bfpsh_ensemble_create": (_I, [ctypes.POINTER(_P), _I, _IP, _I, _I, _I]),
    "bfpsh_ensemble_destroy": (_I, [_P]),
    "bfpsh_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "bfpsh_ensemble_box": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "bfpsh_ensemble_step": (_I, [_P, _DP, _I]),
    "bfpsh_ensemble_time_steps": (_I, [_P, _I, _DP, _I, _DP]),
    "bfpsh_ensemble_step_separate": (_I, [_P, _DP, _I]),
    "bfpsh_ensemble_time_steps_separate": (_I, [_P, _I, _DP, _I, _DP]),
    "bfpsh_ensemble_synchronize": (_I, [_P]),
    "bfpsh_synchronize": (_I, [_P]),
    "bfpsh_get_iteration": (_I, [_P, _IP]),
    "bfpsh_set_iteration": (_I, [_P, _I]),
    "bfpsh_time_steps": (_I, [_P, _I, _D, _I, _DP]),
    "bfpsh_profile_enable": (_I, [_P, _I]),
    "bfpsh_profile_get": (_I, [_P, _I, ctypes.c_char_p, _I, _DP, _IP]),
    "bfpsh_device_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
}

_lib = None


class BfpshError(RuntimeError):
    pass


def load():
    """Load libbfps_hip.so and bind every entry point of include/bfpsh.h."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise BfpshError("%s is not built (run `python -m bfps_amd.build`)" % LIB)
        lib = ctypes.CDLL(LIB)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(ctx, rc, what):
    if rc != 0:
        msg = load().bfpsh_last_error(ctx)
        raise BfpshError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))
