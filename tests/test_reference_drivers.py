"""The reference's own drivers (test_compare/gpe_imag.c, gpe_real.c,
test_vortex/ite_vortex.c, rte_vortex.c) compiled and linked against
include/ and libcugpe_amd/libgpe_hip.so after exactly the source edits the
integration notes list: their three CUDA runtime calls (cudaSetDevice,
cudaHostAlloc, cudaFreeHost) become gpe_set_device / gpe_host_alloc /
gpe_host_free, the CUDA headers go, and cudaError_t becomes int.  Each driver
is copied into a temporary directory at test time, never into the repository,
and built the way compare_with_orginal.sh:73,80 and test_vortex/Makefile
build them (g++, -DNX -DNY -DNZ -DINTERACTIONS, -DDEV for the dev headers).
Skipped where the reference tree is absent (it is not part of this
repository)."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "drivers"))
from ref_drivers import REF, port_driver  # noqa: E402

DRIVERS = ["test_compare/gpe_imag.c", "test_compare/gpe_real.c", "test_vortex/ite_vortex.c",
           "test_vortex/rte_vortex.c"]


def test_port_edits_touch_only_the_cuda_runtime_calls():
    src = ("#include <cuda.h>\n#include <cuda_runtime.h>\n#include <cufft.h>\n#include <math.h>\n"
           "cudaError_t err;\nerr=cudaSetDevice( device );\nif(err != cudaSuccess) return 1;\n"
           "err=cudaHostAlloc((void**) &psi , sizeof(cuCplx)*nxyz, cudaHostAllocDefault );\n"
           "err=cudaHostAlloc( &psi_init , sizeof(cplx)*nxyz, cudaHostAllocDefault );\n"
           "gpe_exec( gpe_evolve_qf(100), ierr );\ncudaFreeHost(psi);\n")
    out = port_driver(src)
    assert out == ("#include \"gpe_engine.h\"\n#include <math.h>\nint err;\nerr=gpe_set_device( device );\n"
                   "if(err != 0) return 1;\n"
                   "err=gpe_host_alloc((void**)&psi, sizeof(cuCplx)*nxyz);\n"
                   "err=gpe_host_alloc((void**)&psi_init, sizeof(cplx)*nxyz);\n"
                   "gpe_exec( gpe_evolve_qf(100), ierr );\ngpe_host_free(psi);\n")


@pytest.mark.parametrize("dev", [True, False], ids=["dev", "orginal"])
@pytest.mark.parametrize("driver", DRIVERS)
def test_reference_driver_builds_against_the_engine(driver, dev, tmp_path):
    """Both spellings of the drivers: -DDEV (the dev/ library) and without it,
    where gpe_imag.c / gpe_real.c typedef cuCplx from cuDoubleComplex
    (test_compare/gpe_imag.c:33-35)."""
    path = os.path.join(REF, driver)
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    cxx = shutil.which("g++")
    lib = os.path.join(ROOT, "libcugpe_amd", "libgpe_hip.so")
    if not cxx or not os.path.exists(lib):
        pytest.skip("no g++ or libgpe_hip.so not built")
    src = port_driver(open(path).read())
    assert "cuda" not in re.sub(r"//.*|/\*.*?\*/", "", src, flags=re.S).replace("cudaSetDevice(%d)", ""), driver
    c = tmp_path / os.path.basename(driver)
    c.write_text(src)
    exe = str(tmp_path / "driver")
    cmd = [cxx, "-std=gnu++14", "-O1", "-fopenmp", "-w", "-DNX=32", "-DNY=32", "-DNZ=32", "-DINTERACTIONS=0",
           "-DVEXT=0"] + (["-DDEV"] if dev else []) + ["-I" + os.path.join(ROOT, "include"), str(c), "-o", exe,
           "-L" + os.path.dirname(lib), "-lgpe_hip", "-Wl,-rpath," + os.path.dirname(lib),
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], stdout=subprocess.PIPE, text=True).stdout
    assert "gpe_create_engine" in nm and "cuda" not in nm
