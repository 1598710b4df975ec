"""Build libbfps_hip.so (gfx950) in-tree with hipcc.

The launch sequences are instantiated once per (precision, N) pair from
csrc/impl_inst.hip; the objects compile in parallel and link with the C ABI
unit csrc/bfpsh.hip into bfps_amd/libbfps_hip.so.

    python -m bfps_amd.build [-j JOBS] [--sizes 32,64,...]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbfps_hip.so")
SIZES = (16, 32, 48, 64, 128, 256, 512, 1024)
ARCH = os.environ.get("BFPSH_ARCH", "gfx950")
# packed FP32 instructions are disabled on purpose (see DESIGN.md notes on gfx950)
FLAGS = ["-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in p or os.path.exists(p):
            return p
    return "hipcc"


def _compile(args):
    src, obj, defs = args
    cmd = [_hipcc()] + FLAGS + defs + ["-c", src, "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, r.returncode, errs, time.time() - t0


def _deps(src, seen=None):
    """The file and the quoted includes it reaches (transitively)."""
    seen = set() if seen is None else seen
    src = os.path.normpath(src)
    if src in seen or not os.path.exists(src):
        return seen
    seen.add(src)
    for line in open(src):
        t = line.strip()
        if t.startswith("#include") and '"' in t:
            _deps(os.path.join(os.path.dirname(src), t.split('"')[1]), seen)
    return seen


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in _deps(src))


def sources():
    out = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))
           if f.endswith((".hip", ".hpp", ".h"))]
    return out


def build(jobs=None, sizes=SIZES, verbose=True, out=None, defines=(), incremental=False):
    """Compile and link; `out` / `defines` build an experimental variant;
    `incremental` recompiles only objects older than their sources."""
    lib = out or LIB
    bdir = BUILD if not defines else BUILD + "_" + "_".join(d.replace("=", "") for d in defines)
    os.makedirs(bdir, exist_ok=True)
    extra = ["-D" + d for d in defines]
    tasks = [(os.path.join(CSRC, "bfpsh.hip"), os.path.join(bdir, "bfpsh.o"), extra),
             (os.path.join(CSRC, "particles.hip"), os.path.join(bdir, "particles.o"), extra)]
    for tn, t in (("f32", "float"), ("f64", "double")):
        for n in sizes:
            tasks.append((os.path.join(CSRC, "impl_inst.hip"),
                          os.path.join(bdir, "impl_%s_%d.o" % (tn, n)),
                          ["-DBFPSH_T=%s" % t, "-DBFPSH_TN=%s" % tn, "-DBFPSH_N=%d" % n] + extra))
    jobs = jobs or min(8, os.cpu_count() or 4)
    t0 = time.time()
    failed = []
    todo = [t for t in tasks if not incremental or _stale(t[0], t[1])]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, rc, errs, dt in ex.map(_compile, todo):
            if verbose:
                print("  %-28s %5.1fs%s" % (os.path.basename(obj), dt, "" if rc == 0 else "  FAILED"))
            if rc != 0:
                failed.append((obj, errs))
            elif errs and verbose:
                print(errs)
    if failed:
        for obj, errs in failed:
            print("---", obj, "\n", errs, file=sys.stderr)
        raise RuntimeError("HIP compilation failed")
    tmp = lib + ".tmp"
    objs = [t[1] for t in tasks]
    subprocess.check_call([_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, lib)
    if verbose:
        print("built %s in %.1fs" % (lib, time.time() - t0))
    return lib


CPP = os.path.join(os.path.dirname(HERE), "tests", "cpp")
DRIVER_SRC = os.path.join(CPP, "nsve_particles.cpp")
DRIVER = os.path.join(CPP, "nsve_particles")
# (source, binary, links the library): the array-fed driver, the file-based
# NSVEparticles job, the HDF5-subset tool
CPP_PROGRAMS = ((DRIVER_SRC, DRIVER, True),
                (os.path.join(CPP, "nsve_particles_job.cpp"), os.path.join(CPP, "nsve_particles_job"), True),
                (os.path.join(CPP, "h5_check.cpp"), os.path.join(CPP, "h5_check"), False))


def build_driver(lib=None, out=None, verbose=True):
    """Compile the C++ programs over include/bfps_hip.hpp (host code only,
    g++): the NSVEparticles drivers against libbfps_hip.so, and the HDF5
    subset tool.  `out` renames the array-fed driver only."""
    lib = lib or LIB
    inc = os.path.join(os.path.dirname(HERE), "include")
    for src, exe, links in CPP_PROGRAMS:
        if out and src == DRIVER_SRC:
            exe = out
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-Wno-stringop-truncation", "-I" + inc,
               src, "-o", exe]
        if links:
            cmd += [lib, "-Wl,-rpath," + os.path.dirname(os.path.abspath(lib))]
        subprocess.check_call(cmd)
        if verbose:
            print("built %s" % exe)
    return out or DRIVER


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("--sizes", default=None)
    ap.add_argument("--out", default=None, help="library path (experimental variants)")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="NAME=VALUE")
    ap.add_argument("-i", "--incremental", action="store_true", help="recompile only stale objects")
    a = ap.parse_args()
    sz = tuple(int(s) for s in a.sizes.split(",")) if a.sizes else SIZES
    lib = build(a.jobs, sz, out=a.out, defines=tuple(a.defines), incremental=a.incremental)
    if not a.out:
        build_driver(lib)
