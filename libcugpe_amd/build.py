"""Builds libcugpe_amd/libgpe_hip.so for gfx950 with hipcc.

One object per (precision, FFT length) for the FFT pass kernels, one per
precision for the other kernels, one for the host engine / C ABI; objects are
compiled in parallel into libcugpe_amd/_build/ and linked into the shared
library in the package directory (in-tree, so it travels with the tree).

The user hook file (the device functions of INTERACTIONS / VEXT 100, in the
reference's gpe_user_defined.h names) is include/gpe_user_hooks.h unless the
environment variable GPE_USER_HOOKS names another file.

    python -m libcugpe_amd.build            # incremental
    python -m libcugpe_amd.build --force    # rebuild everything
    GPE_USER_HOOKS=my_hooks.h python -m libcugpe_amd.build --force
"""
import argparse
import concurrent.futures as cf
import glob
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgpe_hip.so")
ARCH = os.environ.get("GPE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libgpe_hip.so)")


def fft_lengths():
    src = open(os.path.join(CSRC, "gpe_kernels.hpp")).read()
    m = re.search(r"#define GPE_FFT_LENGTHS\(X\)(.*)", src)
    return [int(x) for x in re.findall(r"X\((\d+)\)", m.group(1))]


COMMON = [
    "-std=c++17",
    "-O3",
    "--offload-arch=" + ARCH,
    "-fPIC",
    # packed FP32 instructions are kept out of every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + INCLUDE,
    "-I" + CSRC,
]


def user_hooks():
    """Absolute path of the user hook file compiled into the library."""
    return os.path.abspath(os.environ.get("GPE_USER_HOOKS") or os.path.join(INCLUDE, "gpe_user_hooks.h"))


def jobs():
    hooks = user_hooks()
    if not os.path.exists(hooks):
        raise RuntimeError("user hook file not found: %s" % hooks)
    hdrs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h")) + [hooks]
    out = []
    inst = os.path.join(CSRC, "gpe_fft_inst.hip")
    for real in ("float", "double"):
        for n in fft_lengths():
            obj = os.path.join(BUILD, "fft_%s_%d.o" % (real, n))
            out.append((obj, [inst], ["-DGPE_REAL=" + real, "-DGPE_FFT_N=%d" % n], hdrs))
    for name in ("gpe_kernels_f32", "gpe_kernels_f64", "gpe_adjoint", "gpe_dipolar", "gpe_engine", "gpe_noalloc", "gpe_fused", "gpe_plane"):
        src = os.path.join(CSRC, name + ".hip")
        out.append((os.path.join(BUILD, name + ".o"), [src], [], hdrs))
    return out


def stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def compile_one(cc, obj, srcs, defs):
    tmp = obj + ".tmp.o"
    cmd = [cc] + COMMON + ['-DGPE_USER_HOOKS_FILE="%s"' % user_hooks()] + defs + ["-c"] + srcs + ["-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (obj, " ".join(cmd), r.stdout))
    os.replace(tmp, obj)
    return obj


def build(force=False, verbose=True, max_jobs=None):
    cc = hipcc()
    os.makedirs(BUILD, exist_ok=True)
    todo = [(o, s, d) for (o, s, d, h) in jobs() if force or stale(o, s + h)]
    n = max_jobs or int(os.environ.get("MAX_JOBS", "0")) or min(16, os.cpu_count() or 4)
    if todo:
        if verbose:
            print("[libcugpe_amd] compiling %d objects with %d jobs" % (len(todo), n), flush=True)
        with cf.ThreadPoolExecutor(max_workers=n) as ex:
            futs = [ex.submit(compile_one, cc, o, s, d) for (o, s, d) in todo]
            for f in cf.as_completed(futs):
                f.result()
    objs = [o for (o, _, _, _) in jobs()]
    if force or todo or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s" % r.stdout)
        os.replace(tmp, LIB)
        if verbose:
            print("[libcugpe_amd] linked %s" % LIB, flush=True)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    build(force=a.force, max_jobs=a.j)


if __name__ == "__main__":
    sys.exit(main())
