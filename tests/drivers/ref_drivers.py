"""Builds the reference's own drivers against this engine, for the drop-in
tests: test_compare/gpe_imag.c and gpe_real.c (run on the GPU by
tests/test_gpu_reference_drivers.py) and test_vortex/rte_vortex.c (run on
the GPU from a .dftc file), test_vortex/ite_vortex.c (build only).  Each driver's source is read from the reference
tree, edited exactly as the integration notes list (its three CUDA runtime
calls become gpe_set_device / gpe_host_alloc / gpe_host_free, the CUDA
headers go, cudaError_t becomes int), written to a temporary directory and
compiled with the reference's own flags (compare_with_orginal.sh:73,80:
g++, -DNX -DNY -DNZ, -DDEV).  Binaries go to tests/drivers/bin/ (not part of
the repository); no reference source is kept.

    python tests/drivers/ref_drivers.py        # when /root/reference exists
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
BIN = os.path.join(HERE, "bin")
LIB = os.path.join(ROOT, "libcugpe_amd", "libgpe_hip.so")
# the lattice of the reference's differential test (compare_with_orginal.sh:9-11)
LATTICE = (256, 32, 32)
# drivers the GPU tests run: (source, lattice)
RUN = {"gpe_imag_dev": ("test_compare/gpe_imag.c", LATTICE), "gpe_real_dev": ("test_compare/gpe_real.c", LATTICE),
       "rte_vortex_dev": ("test_vortex/rte_vortex.c", (32, 32, 32))}


def port_driver(src):
    """The integration notes' edits, applied to a driver's source text: the
    first CUDA header include becomes the engine header (which defines the
    complex types the drivers name before their own include of it), the other
    CUDA header includes go."""
    cuda_inc = r"^[ \t]*#include[ \t]*<(?:cuda|cuda_runtime|cufft)\.h>[ \t]*\n"
    s = re.sub(cuda_inc, '#include "gpe_engine.h"\n', src, count=1, flags=re.M)
    s = re.sub(cuda_inc, "", s, flags=re.M)
    s = re.sub(r"\bcudaError(?:_t)?(\s+err\s*;)", r"int\1", s)
    s = s.replace("cudaSetDevice(", "gpe_set_device(")
    s = re.sub(r"cudaHostAlloc\(\s*(?:\(void\s*\*\*\)\s*)?(&\s*\w+)\s*,\s*([^,]+?)\s*,\s*cudaHostAllocDefault\s*\)",
               r"gpe_host_alloc((void**)\1, \2)", s)
    s = s.replace("cudaFreeHost(", "gpe_host_free(")
    s = s.replace("cudaSuccess", "0")
    return s


def compile_driver(rel_src, exe, lattice=LATTICE, workdir=None, dev=True):
    """Port and compile one driver (dev: with the reference's -DDEV, else the
    orginal spelling); returns (returncode, compiler output)."""
    cxx = shutil.which("g++")
    if not cxx:
        return 1, "no g++"
    src = port_driver(open(os.path.join(REF, rel_src)).read())
    tmp = workdir or tempfile.mkdtemp()
    c = os.path.join(tmp, os.path.basename(rel_src))
    with open(c, "w") as f:
        f.write(src)
    nx, ny, nz = lattice
    cmd = [cxx, "-std=gnu++14", "-O2", "-fopenmp", "-w", "-DNX=%d" % nx, "-DNY=%d" % ny, "-DNZ=%d" % nz,
           "-DINTERACTIONS=0", "-DVEXT=0"] + (["-DDEV"] if dev else []) + ["-I" + os.path.join(ROOT, "include"), c, "-o", exe,
           "-L" + os.path.dirname(LIB), "-lgpe_hip", "-Wl,-rpath,$ORIGIN/../../../libcugpe_amd",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if not workdir:
        shutil.rmtree(tmp, ignore_errors=True)
    return r.returncode, r.stdout


def build():
    """Builds the drivers of RUN into tests/drivers/bin/; a no-op without the reference tree."""
    if not os.path.isdir(REF) or not os.path.exists(LIB):
        return []
    os.makedirs(BIN, exist_ok=True)
    built = []
    for name, (rel_src, lattice) in RUN.items():
        rc, out = compile_driver(rel_src, os.path.join(BIN, name), lattice)
        if rc != 0:
            raise RuntimeError("building %s failed:\n%s" % (rel_src, out))
        built.append(name)
    return built


if __name__ == "__main__":
    print(build())
    sys.exit(0)
