"""Build the gfx950 HIP library libpadeops_hip.so in-tree with hipcc.

Every source under padeops_amd/csrc is compiled (host C++ and HIP alike) by
hipcc for --offload-arch=gfx950 and linked into padeops_amd/libpadeops_hip.so.
When amdflang is present, the Fortran ISO_C_BINDING shim
(padeops_amd/fortran/padeops_hip_mod.F90) and its driver test_shim are built
too, against that library.
Usage:  python -m padeops_amd.build   (or padeops_amd.build.build()).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpadeops_hip.so")
ARCH = os.environ.get("PDO_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
            # keep the compiler from emitting packed FP32 instructions
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(HERE, "..", "include"), "-I", CSRC]


# per-file extras: the MLP kernels take their MFMA results in VGPRs (no
# accumulator-register copies on the way to the GELU epilogue)
FILE_FLAGS = {"pdo_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc()] + flags() + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


FORT = os.path.join(HERE, "fortran")


def amdflang():
    for c in (os.environ.get("PDO_FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose=False):
    """Compile the shim module and link the driver padeops_amd/fortran/test_shim
    (rpath to the library one directory up).  Returns its path, or None when
    no Fortran compiler is installed."""
    fc = amdflang()
    if fc is None:
        return None
    work = os.path.join(OBJ, "fortran")
    os.makedirs(work, exist_ok=True)
    dec = os.path.join(FORT, "padeops_hip_decomp.F90")
    mod = os.path.join(FORT, "padeops_hip_mod.F90")
    drv = os.path.join(FORT, "test_shim.F90")
    exe = os.path.join(FORT, "test_shim")
    steps = [
        [fc, "-O2", "-module-dir", work, "-c", dec, "-o", os.path.join(work, "dec.o")],
        [fc, "-O2", "-module-dir", work, "-c", mod, "-o", os.path.join(work, "mod.o")],
        [fc, "-O2", "-module-dir", work, "-c", drv, "-o", os.path.join(work, "drv.o")],
        [fc, os.path.join(work, "drv.o"), os.path.join(work, "mod.o"), os.path.join(work, "dec.o"), "-L", HERE,
         "-lpadeops_hip",
         "-Wl,-rpath,$ORIGIN/..", "-o", exe],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("Fortran build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", exe)
    return exe


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))
    # the heaviest translation units first
    srcs.sort(key=lambda p: -os.path.getsize(p))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
