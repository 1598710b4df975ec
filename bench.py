"""Benchmark: bfps vorticity-equation step + tracer loop on one MI355X.

Workloads (synthetic, seeded; E(k) ~ k^4 exp(-2 (k/k0)^2) initial field of
total energy 1/2, uniform random tracers, AB4, dt = 0.5/N):

  512   (default) precision form: 512^3, FP32 FFTs / nonlinear arithmetic with
        FP64 integrating factors, reductions and tracer positions, nu = 0.0012,
        linear band forcing 0.5 on 2 <= |k| <= 4, 2^22 tracers, 6-neighbour
        smoothness-2 interpolation of the velocity; velocity-gradient
        sampling timed separately.
  1024  large single box: 1024^3 FP32, nu = 4e-4, fixed energy-injection rate
        0.1 on 1.5 <= |k| <= 3.5, 2^24 tracers, 4-neighbour smoothness-1.

One step = NSVEparticles::step (tracer completeLoop with u(t), then the
IF-SSP-RK3 vorticity step).  The same workload in FP64 is timed beside it,
and the accuracy leg runs FP32 and FP64 from the same seed for --accuracy
steps and reports the FP32 error (vorticity L2, per-shell energy spectrum,
tracer positions) against FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 512|1024]
                    [--accuracy STEPS] [--ensemble-steps K] [--no-cpu] [--no-fp64]

The ensemble leg (96 boxes of 32^3..128^3, batched vs box by box) is
reported under "ensemble".

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
LDS_PEAK = 128 * 256 * 2.4e9  # B/s: 128 B/clk per CU, 256 CUs, 2.4 GHz
FP64_PEAK = 78.6e12  # MI355X FP64 vector, FLOP/s

WORKLOADS = {
    "512": dict(n=512, nu=0.0012, forcing="linear", famp=0.5, fk0=2.0, fk1=4.0, np=2 ** 22, nb=6, m=2, k0=4.0,
                name="512^3 precision form: FP32 vorticity step + 2^22 tracers n6 m2 AB4"),
    "1024": dict(n=1024, nu=4e-4, forcing="fixed_energy_injection_rate", famp=0.1, fk0=1.5, fk1=3.5, np=2 ** 24,
                 nb=4, m=1, k0=4.0,
                 name="1024^3 FP32 vorticity step, fixed energy injection, + 2^24 tracers n4 m1 AB4"),
}


def step_bytes(n, es, with_tracer_u):
    """Algorithmic HBM bytes of one fused vorticity step: 141 field volumes
    (N^3 elements of size es) + 3 for the real velocity handed to the
    tracers."""
    return (141 + (3 if with_tracer_u else 0)) * n ** 3 * es


# algorithmic bytes per launch of each kernel, in field volumes N^3 * es
# ("inv_z_bs" is reported as one kernel: the kx tile holding kx = 0 runs as a
# separate launch "inv_z_bs_t0" with the symmetrize mirror, merged below)
KERNEL_VOLUMES = {
    "inv_z_bs": 9,         # read omega (3 planes), write u, omega after the z pass (6)
    "inv_y": 12,           # 6 planes read + written
    "rows_cross": 9,       # read 6, write 3 (+3 real velocity in stage 1)
    "fwd_z": 6,            # 3 planes read + written
    "fwd_y_update1": 9,    # read N (3), v0 (3), write v1 (3)
    "fwd_y_update2": 12,   # read N, v0, v1, write v2
    "fwd_y_update3": 12,   # read N, v0, v2, write v0
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_%d_fp32.txt")


def merge_profile(prof):
    """Fold the kx-tile-0 launch of the Biot-Savart pass into "inv_z_bs";
    the tracer velocity loop is "interp_ab" whichever kernel form ran (the
    column walk and the global gather are profiled as "interp_ab_col" and
    "interp_ab_gather")."""
    out = dict(prof)
    if "inv_z_bs_t0" in out and "inv_z_bs" in out:
        t0_ms, _ = out.pop("inv_z_bs_t0")
        ms, cnt = out["inv_z_bs"]
        out["inv_z_bs"] = (ms + t0_ms, cnt)
    for form in ("interp_ab_col", "interp_ab_gather"):
        if form in out:
            ms, cnt = out.pop(form)
            a, b = out.get("interp_ab", (0.0, 0))
            out["interp_ab"] = (ms + a, cnt + b)
    return out


def tracer_form(prof):
    """Kernel form of the tracer velocity loop in a raw profile."""
    if "interp_ab_col" in prof:
        return "column walk"
    if "interp_ab_gather" in prof:
        return "global gather"
    return "bricks" if "interp_ab" in prof else None


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE x calibration + WRITE_SIZE), or None."""
    path = TRAFFIC_FILE % n
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 2 and f[0] == kernel:
            return float(f[1])
    return None


def tracer_bytes(np_, nsteps=4):
    """positions read+write (48 B), history read (8 (nsteps-1) 3 B) and
    write (24 B) per tracer."""
    return np_ * (48 + 24 * (nsteps - 1) + 24)


def make_case(B, ic, wl, precision, seed=1):
    import numpy as np
    n = wl["n"]
    fs = B.vorticity_equation("bench", n, precision=precision)
    fs.nu, fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = wl["nu"], wl["famp"], wl["fk0"], wl["fk1"], wl["forcing"]
    w = ic.solenoidal_vorticity(n, seed=seed, k0=wl["k0"], dtype=np.complex64 if precision == "single" else complex)
    fs.io_checkpoint(read=True, data=w)
    del w
    ps = None
    if wl["np"]:
        ps = B.particles_system(fs, wl["np"], wl["nb"], wl["m"], 4, step_idx=1)
        ps.put(ic.uniform_positions(wl["np"], seed=seed + 1))
        # one untimed sample settles the tracer kernel form (chosen by timing
        # at MEASURE rigor) before any timed region, whatever --warmup is
        fs.velocity_rspace()
        ps.sample("velocity")
    return fs, ps


def run_precision(B, ic, wl, precision, steps, warmup, gradient=True):
    fs, ps = make_case(B, ic, wl, precision)
    n = wl["n"]
    dt = 0.5 / n
    fs._push_physics()
    for _ in range(warmup):
        B.NSVEparticles_step(fs, dt)
    fs.synchronize()
    fs.profile(True)
    t0 = time.perf_counter()
    ms = fs.time_steps(steps, dt, with_particles=True)
    wall = time.perf_counter() - t0
    prof = fs.profile_report()
    fs.profile(False)
    grad_s = None
    if gradient and ps is not None:
        # velocity gradient (9 components) sampled at all tracers
        fs.synchronize()
        t1 = time.perf_counter()
        reps = 2
        for _ in range(reps):
            fs.gradient_rspace()
            ps.sample("gradient")
        fs.synchronize()
        grad_s = (time.perf_counter() - t1) / reps
    stats = None
    if gradient and ps is not None:
        stats = stats_leg(B, fs, ps, n, 4 if precision == "single" else 8)
    res = {"ms_per_step": ms / steps, "wall_ms_per_step": 1e3 * wall / steps, "profile": prof,
           "grad_sample_s": grad_s, "device_bytes": fs.device_bytes(), "stats": stats,
           "pass_shapes": {k: {"values_per_thread": v[0], "columns": v[1], "persistent": bool(v[4])}
                           for k, v in fs.pass_shapes().items()}}
    fs.close()
    return res


def stats_leg(B, fs, ps, n, es):
    """One NSVEparticles::do_stats (bfps/cpp/full_code/NSVEparticles.cpp:70-98)
    on the device: vorticity and velocity statistics (cospectrum + real-space
    moments and histograms of 256 bins), then the velocity and the Lagrangian
    acceleration sampled at every tracer.  The real-space statistics kernel
    reads its 3 real planes once: its roofline is 3 N^3 e bytes per launch."""
    B.NSVEparticles_do_stats(fs, ps, 1.0, 1.0, 256)  # warm-up (allocations)
    fs.synchronize()
    fs.profile(True)
    t0 = time.perf_counter()
    B.NSVEparticles_do_stats(fs, ps, 1.0, 1.0, 256)
    fs.synchronize()
    wall = time.perf_counter() - t0
    prof = fs.profile_report()
    fs.profile(False)
    out = {"do_stats_ms": 1e3 * wall, "note": "wall time incl. the host copies of the sampled tracer values"}
    if "rspace_stats" in prof:
        ms, cnt = prof["rspace_stats"]
        avg = ms / cnt
        byt = 3 * n ** 3 * es
        out["rspace_stats"] = {"avg_ms": avg, "bytes_per_launch": byt, "achieved_gbps": byt / (avg / 1e3) / 1e9,
                               "frac": byt / (avg / 1e3) / HBM_PEAK}
    for k in ("pressure", "pressure_uu", "lagrangian_acceleration"):
        if k in prof:
            out[k + "_ms"] = prof[k][0] / prof[k][1]
    return out


def roofline(prof, n, es, np_, nb, traffic_ok, steps):
    """Roofline of the kernel with the largest share of the step: achieved =
    algorithmic bytes per launch / average launch duration (HIP events on
    the context stream)."""
    form = tracer_form(prof)
    prof = merge_profile(prof)
    name, (ms, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = ms / cnt / 1e3
    if name in KERNEL_VOLUMES:
        bytes_per_launch = KERNEL_VOLUMES[name] * n ** 3 * es
        note = "%g field volumes of %d^3 x %d B" % (KERNEL_VOLUMES[name], n, es)
    elif name == "interp_ab":
        # tracer state (144 B per tracer) + one read of the 3 real velocity volumes
        bytes_per_launch = tracer_bytes(np_) + 3 * n ** 3 * es
        note = ("tracer state 144 B x %d + 3 real velocity volumes; the kernel also reads %d stencil "
                "values per tracer from LDS" % (np_, 3 * (2 * nb + 2) ** 3))
    else:
        return {"kernel": name, "avg_ms": ms / cnt}
    ach = bytes_per_launch / avg_s
    traffic = measured_traffic(name, n) if traffic_ok else None
    out = {"bound": "hbm", "kernel": name, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": ach / HBM_PEAK, "traffic": traffic, "avg_ms": ms / cnt, "bytes_per_launch": bytes_per_launch,
           "bytes_note": note, "launches_per_step": cnt / steps}
    if name == "interp_ab":
        # LDS traffic of the tracer kernel: every tracer reads its (2n+2)^3
        # stencil of each component from LDS; the windows are written once
        # per component: column walk, each y plane of a 32 x 16 (x, z)
        # column with its halo (x rows widened to aligned 16-byte groups)
        # once per walk; bricks, each brick's tile with its halo of 2n+1
        # cells on all three axes
        w = 2 * nb + 2
        xa = 16 // es
        xs = (xa - nb % xa) % xa
        if form == "column walk":
            kx, kz = 32, 16
            wxt = (kx + w - 1 + xs + xa - 1) // xa * xa
            windows = 3 * (n // kx) * (n // kz) * (n + w - 1) * wxt * (kz + w - 1)
        else:
            kx, ky, kz = (16, 8, 16 if w <= 10 else 8) if es == 4 else (8, 8, 8)
            wxt = (kx + w - 1 + xs + xa - 1) // xa * xa
            windows = 3 * (n ** 3 // (kx * ky * kz)) * wxt * (ky + w - 1) * (kz + w - 1)
        lds_bytes = np_ * 3 * w ** 3 * es + windows * es
        out["form"] = form
        out["lds"] = {"bytes_per_launch": lds_bytes, "achieved": lds_bytes / avg_s / 1e9, "peak": LDS_PEAK / 1e9,
                      "unit": "GB/s", "frac": lds_bytes / avg_s / LDS_PEAK}
        # FP64 arithmetic of the reference's sum
        # (particles_field_computer.hpp:142-150): per stencil point the
        # weight product bz by bx (2 flops) and, per component, value x
        # weight accumulated (2 flops); the kernels do it factorized (fewer
        # flops) but convert every value to FP64 first, and their bound is
        # FP64 VALU issue
        flops = np_ * w ** 3 * (2 + 3 * 2)
        out["fp64"] = {"flops_per_launch": flops, "achieved": flops / avg_s / 1e12, "peak": FP64_PEAK / 1e12,
                       "unit": "TFLOP/s", "frac": flops / avg_s / FP64_PEAK}
    return out


def progress(msg):
    """One line per finished leg on stderr (stdout carries only the JSON line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def accuracy(B, ic, wl, steps):
    """FP32 against FP64 from the same seed after `steps` NSVEparticles steps:
    relative L2 of the vorticity (half spectrum, nxmodes weights = real-space
    L2), per-shell energy spectrum error, tracer position drift."""
    import numpy as np
    n = wl["n"]
    dt = 0.5 / n
    out = {}
    for prec in ("double", "single"):
        fs, ps = make_case(B, ic, wl, prec)
        for _ in range(steps):
            B.NSVEparticles_step(fs, dt)
        progress("accuracy leg: %d %s steps done" % (steps, prec))
        w = fs.cvorticity.astype(np.complex128)
        spec = fs.cospectrum("velocity", "velocity")
        pos = ps.get()[0] if ps is not None else None
        out[prec] = (w, np.trace(spec, axis1=1, axis2=2) / 2, pos)
        fs.close()
        del fs, ps
    (w64, e64, p64), (w32, e32, p32) = out["double"], out["single"]
    wt = np.ones(w64.shape[2])
    wt[1:] = 2.0
    d = np.abs(w32 - w64) ** 2
    num = float(np.sqrt(np.einsum("abkc,k->", d, wt)))
    den = float(np.sqrt(np.einsum("abkc,k->", np.abs(w64) ** 2, wt)))
    del d
    ok = e64 > 1e-12 * e64.max()
    shell_err = np.abs(e32[ok] - e64[ok]) / e64[ok]
    res = {"steps": steps, "reference": "the same workload in FP64 on the GPU (checked against the CPU oracle at "
                                        "256^3 and 512^3, and at 1024^3 on the embedded low-mode step, "
                                        "by the parity tests)",
           "vorticity_rel_l2": num / den,
           "energy_spectrum_rel_err_max": float(shell_err.max()),
           "energy_spectrum_rel_err_median": float(np.median(shell_err))}
    if p64 is not None:
        dp = np.linalg.norm(p32 - p64, axis=1)
        res["tracer_drift_max"] = float(dp.max())
        res["tracer_drift_rms"] = float(np.sqrt(np.mean(dp ** 2)))
        res["grid_spacing"] = 2 * np.pi / n
    return res


ENSEMBLE = [(32, 32), (32, 48), (24, 64), (8, 128)]  # (boxes, N)


def ensemble_bench(B, ic, steps, warmup=2, seed=5):
    """96 independent FP32 boxes (32 x 32^3, 32 x 48^3, 24 x 64^3, 8 x 128^3),
    each with its own seed, nu ~ U[0.01, 0.05], k0 in {2, 3, 4}, linear band
    forcing, dt = 0.5/N, and log-uniform [1e3, 1e5] tracers (n1 m1 AB4, the
    reference's defaults).  The batched form (one graph launch of batched field
    kernels per step + the tracer loop of each box) against the same step issued
    box by box."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [n for cnt, n in ENSEMBLE for _ in range(cnt)]
    ens = B.ensemble(sizes, precision="single")
    total_np = 0
    parts = []
    for i, n in enumerate(sizes):
        fs = ens.boxes[i]
        fs.nu = float(rng.uniform(0.01, 0.05))
        fs.famplitude, fs.fk0, fs.fk1, fs.forcing_type = 0.5, 2.0, 4.0, "linear"
        fs.io_checkpoint(read=True, data=ic.solenoidal_vorticity(n, seed=seed * 1000 + i,
                                                                 k0=float(rng.choice([2.0, 3.0, 4.0]))))
        npi = int(round(10 ** rng.uniform(3, 5)))
        total_np += npi
        ps = B.particles_system(fs, npi, 1, 1, 4, step_idx=1)
        ps.put(ic.uniform_positions(npi, seed=seed * 1000 + 500 + i))
        parts.append(ps)
    dt = np.array([0.5 / n for n in sizes])
    for _ in range(warmup):
        ens.step(dt)
    ens.synchronize()
    ms_b = ens.time_steps(steps, dt)
    for _ in range(warmup):
        ens.step_separate(dt)
    ens.synchronize()
    ms_s = ens.time_steps_separate(steps, dt)
    points = sum(n ** 3 for n in sizes)
    ens.close()
    return {"boxes": len(sizes), "grid_points": points, "tracers": total_np, "steps": steps,
            "batched_ms_per_step": ms_b / steps, "per_box_ms_per_step": ms_s / steps,
            "batched_boxes_steps_per_s": len(sizes) * steps / (ms_b / 1e3),
            "per_box_boxes_steps_per_s": len(sizes) * steps / (ms_s / 1e3),
            "batched_grid_pt_steps_per_s": points * steps / (ms_b / 1e3),
            "speedup": ms_s / ms_b}


def cpu_info():
    """CPU model and logical-core counts of this host (nproc: the cores this
    process may run on)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"model": model, "nproc": nproc, "cpu_count": os.cpu_count()}


def cpu_threads_default():
    """All the cores this process may use, within the OpenMP thread limit
    the host sets (OMP_NUM_THREADS), which the CPU restatement obeys."""
    nproc = cpu_info()["nproc"]
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(nproc, int(omp)) if omp.isdigit() and int(omp) > 0 else nproc


def cpu_baseline(ic, wl, threads, single_n=256):
    """The oracle (CPU restatement, C++/OpenMP) on `threads` host threads, on a
    sample of the same workload: after one untimed warm-up step (first touch
    of the buffers, twiddle tables), one NSVEparticles step at full size --
    the velocity transform, the tracer loop over all tracers, the FP32
    vorticity step.  Then one FP32 vorticity step of a `single_n`^3 box on a
    single thread, for per-core throughput."""
    from oracle import oracle as O
    n = wl["n"]
    O.build()
    O.set_threads(threads)
    s = O.Solver(n, precision="single")
    s.set_physics(wl["nu"], wl["forcing"], wl["famp"], wl["fk0"], wl["fk1"])
    s.set_cfield("vorticity", ic.solenoidal_vorticity(n, seed=1, k0=wl["k0"]))
    s.io_projection()
    dt = 0.5 / n
    p = None
    if wl["np"]:
        p = O.Particles(s, wl["np"], wl["nb"], wl["m"], 4, step_idx=1)
        p.set_state(ic.uniform_positions(wl["np"], seed=2))
    t0 = time.perf_counter()
    O.nsve_particles_step(s, p, dt) if p is not None else s.step(dt)
    t_warm = time.perf_counter() - t0
    progress("CPU baseline: warm-up step %.1f s" % t_warm)
    t0 = time.perf_counter()
    s.velocity_rspace()
    t_vel = time.perf_counter() - t0
    t_tr = 0.0
    if p is not None:
        t0 = time.perf_counter()
        p.complete_loop(dt)
        t_tr = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.step(dt)
    t_step = time.perf_counter() - t0
    del p, s
    total = t_vel + t_tr + t_step
    progress("CPU baseline: timed step %.1f s on %d threads" % (total, threads))
    # one core
    O.set_threads(1)
    s1 = O.Solver(single_n, precision="single")
    s1.set_physics(wl["nu"], wl["forcing"], wl["famp"], wl["fk0"], wl["fk1"])
    s1.set_cfield("vorticity", ic.solenoidal_vorticity(single_n, seed=1, k0=wl["k0"]))
    s1.io_projection()
    s1.step(0.5 / single_n)
    t0 = time.perf_counter()
    s1.step(0.5 / single_n)
    t_one = time.perf_counter() - t0
    del s1
    O.set_threads(threads)
    return {"value": n ** 3 / total, "unit": "grid-pt*steps/s", "cores": threads, "kind": "port",
            "sample": ("one NSVEparticles step of the %d^3 FP32 workload with all %d tracers by the CPU oracle "
                       "(after one untimed warm-up step): velocity transform %.1f s, tracer loop %.1f s, "
                       "vorticity step %.1f s" % (n, wl["np"], t_vel, t_tr, t_step)),
            "cpu": cpu_info(),
            "particle_interps_per_s": (wl["np"] / t_tr) if t_tr else None,
            "vorticity_step_s": t_step,
            "single_core": {"n": single_n, "vorticity_step_s": t_one,
                            "grid_pt_steps_per_s_per_core": single_n ** 3 / t_one,
                            "note": "FP32 vorticity step (no tracers) of a %d^3 box on one thread" % single_n}}


def rocfft_leg(n, reps=5):
    """rocFFT, the vendor FFT library, timed as a reference only: the
    3-component FP32 r2c and c2r of an n^3 box (batched 3D transforms, not
    in place, default plans, on one stream), as one vorticity step needs 6
    vector c2r and 3 vector r2c (2 inverse and 1 forward per RK stage).
    The library is loaded through ctypes; nothing of this build calls it."""
    import ctypes
    try:
        rf = ctypes.CDLL("librocfft.so")
    except OSError as e:
        return {"error": "librocfft not loadable: %s" % e}
    hip = ctypes.CDLL("libamdhip64.so")
    P, S = ctypes.c_void_p, ctypes.c_size_t
    rf.rocfft_setup()
    ver = ctypes.create_string_buffer(64)
    rf.rocfft_get_version_string(ver, S(64))
    stream = P()
    hip.hipStreamCreate(ctypes.byref(stream))
    nr, nc = n ** 3, n * n * (n // 2 + 1)
    rbuf, cbuf = P(), P()
    if hip.hipMalloc(ctypes.byref(rbuf), S(3 * nr * 4)) or hip.hipMalloc(ctypes.byref(cbuf), S(3 * nc * 8)):
        return {"error": "hipMalloc failed"}
    hip.hipMemset(rbuf, 0, S(3 * nr * 4))
    lengths = (S * 3)(n, n, n)
    out = {"version": ver.value.decode(), "n": n, "components": 3, "precision": "f32"}
    for name, ttype, src, dst in (("r2c", 2, rbuf, cbuf), ("c2r", 3, cbuf, rbuf)):
        plan, info, work = P(), P(), P()
        if rf.rocfft_plan_create(ctypes.byref(plan), 1, ttype, 0, S(3), lengths, S(3), None):
            return {"error": "rocfft_plan_create failed (%s)" % name}
        wsz = S(0)
        rf.rocfft_plan_get_work_buffer_size(plan, ctypes.byref(wsz))
        rf.rocfft_execution_info_create(ctypes.byref(info))
        rf.rocfft_execution_info_set_stream(info, stream)
        if wsz.value:
            hip.hipMalloc(ctypes.byref(work), wsz)
            rf.rocfft_execution_info_set_work_buffer(info, work, wsz)
        ins, outs = (P * 1)(src), (P * 1)(dst)
        for _ in range(2):
            rf.rocfft_execute(plan, ins, outs, info)
        a, b = P(), P()
        hip.hipEventCreate(ctypes.byref(a))
        hip.hipEventCreate(ctypes.byref(b))
        hip.hipEventRecord(a, stream)
        for _ in range(reps):
            rf.rocfft_execute(plan, ins, outs, info)
        hip.hipEventRecord(b, stream)
        hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), a, b)
        t = ms.value / reps
        byt = 3 * (nr * 4 + nc * 8)
        out[name] = {"ms": t, "bytes": byt, "achieved_gbps": byt / (t / 1e3) / 1e9,
                     "frac": byt / (t / 1e3) / HBM_PEAK, "work_buffer_bytes": wsz.value}
        hip.hipEventDestroy(a)
        hip.hipEventDestroy(b)
        rf.rocfft_execution_info_destroy(info)
        rf.rocfft_plan_destroy(plan)
        if wsz.value:
            hip.hipFree(work)
    hip.hipFree(rbuf)
    hip.hipFree(cbuf)
    hip.hipStreamDestroy(stream)
    out["transforms_per_step_ms"] = 6 * out["c2r"]["ms"] + 3 * out["r2c"]["ms"]
    out["note"] = ("6 vector c2r + 3 vector r2c: the transforms of one vorticity step alone, without its "
                   "pointwise work, the fused Biot-Savart / cross product / update or the tracer velocity")
    return out


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size: rocprofv3's FETCH_SIZE /
    WRITE_SIZE of this copy calibrate the counters of the other kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="512")
    ap.add_argument("--np", type=int, default=None, help="override the tracer count")
    ap.add_argument("--accuracy", type=int, default=100, help="steps of the FP32-vs-FP64 leg (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=20,
                    help="steps of the 96-box ensemble leg (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rocfft", action="store_true", help="skip the rocFFT reference leg")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="finish with a device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import bfps_amd as B
    from bfps_amd import ic

    wl = dict(WORKLOADS[args.workload])
    default_case = args.np is None  # the tracer counts the committed PMC summaries were measured with
    if args.np is not None:
        wl["np"] = args.np
    n, np_ = wl["n"], wl["np"]
    r32 = run_precision(B, ic, wl, "single", args.steps, args.warmup)
    ms = r32["ms_per_step"]
    progress("FP32 leg: %.3f ms per step" % ms)
    tb = tracer_bytes(np_)
    alg = step_bytes(n, 4, np_ > 0) + tb
    out = {
        "metric": "grid-pt*steps/s + particle interps/s, FP32 vs FP64 at 512^3, % HBM peak",
        "value": n ** 3 / (ms / 1e3),
        "unit": "grid-pt*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision_note": "f32 field storage, FFTs and cross product; f64 integrating factors, forcing "
                          "amplitudes, reductions, interpolation weights and stencil sums (every f32 field value "
                          "is converted to f64 before it is accumulated), tracer positions and history",
        "data": "synthetic: seeded solenoidal field E(k)~k^4 exp(-2(k/%g)^2), seeded uniform tracers" % wl["k0"],
        "config": {"workload": wl["name"], "n": n, "nu": wl["nu"], "dt": 0.5 / n, "forcing": wl["forcing"],
                   "tracers": np_, "neighbours": wl["nb"], "smoothness": wl["m"]},
        "particle_interps_per_s": np_ / (ms / 1e3),
        "gradient_sample_interps_per_s": np_ / r32["grad_sample_s"] if r32["grad_sample_s"] else None,
        "stats": r32["stats"],
        "hbm_frac_step": alg / (ms / 1e3) / HBM_PEAK,
        "algorithmic_bytes_per_step": alg,
        "roofline": roofline(r32["profile"], n, 4, np_, wl["nb"], default_case, args.steps),
        "kernel_ms": {k: v[0] / v[1] for k, v in merge_profile(r32["profile"]).items()},
        "wall_ms_per_step": r32["wall_ms_per_step"],
        "device_gb": r32["device_bytes"] / 1e9,
        "pass_shapes": r32["pass_shapes"],
    }
    kp = merge_profile(r32["profile"])
    if np_ > 0 and "interp_ab" in kp:
        # the tracer kernel alone: interpolations (one tracer x one 3-component
        # field) and stencil-point x component products per second
        t_ia = kp["interp_ab"][0] / kp["interp_ab"][1] / 1e3
        w = 2 * wl["nb"] + 2
        out["tracer_kernel"] = {"form": tracer_form(r32["profile"]), "ms": t_ia * 1e3, "interps_per_s": np_ / t_ia,
                                "stencil_point_components_per_s": np_ * w ** 3 * 3 / t_ia}
    if not args.no_fp64:
        s64 = max(2, args.steps // 2)
        r64 = run_precision(B, ic, wl, "double", s64, 1, gradient=False)
        ms64 = r64["ms_per_step"]
        progress("FP64 leg: %.3f ms per step" % ms64)
        out["fp64"] = {"value": n ** 3 / (ms64 / 1e3), "ms_per_step": ms64,
                       "particle_interps_per_s": np_ / (ms64 / 1e3),
                       "hbm_frac_step": (step_bytes(n, 8, np_ > 0) + tb) / (ms64 / 1e3) / HBM_PEAK,
                       "roofline": roofline(r64["profile"], n, 8, np_, wl["nb"], False, s64),
                       "tracer_form": tracer_form(r64["profile"]),
                       "kernel_ms": {k: v[0] / v[1] for k, v in merge_profile(r64["profile"]).items()}}
    if args.accuracy > 0 and not args.no_fp64:
        out["accuracy"] = accuracy(B, ic, wl, args.accuracy)
        progress("accuracy leg done")
    if args.ensemble_steps > 0:
        out["ensemble"] = ensemble_bench(B, ic, args.ensemble_steps)
        progress("ensemble leg done")
    if not args.no_rocfft:
        out["rocfft_reference"] = rocfft_leg(n)
        progress("rocFFT reference leg done")
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_threads_default()
        out["cpu_baseline"] = cpu_baseline(ic, wl, threads)
        progress("CPU baseline done")
    else:
        out["cpu_baseline"] = None
    if args.copy_gib > 0:
        copy_calibration(int(args.copy_gib * 2 ** 30))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
