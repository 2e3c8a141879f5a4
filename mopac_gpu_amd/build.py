"""Build libmopachip.so (gfx950) in-tree with hipcc.

Usage: python -m mopac_gpu_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmopachip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("MHIP_FC", "/opt/rocm/bin/amdflang")
FORTRAN_MODULE = os.path.join(HERE, "fortran", "mopac_hip.F90")
FORTRAN_TEST = os.path.join(ROOT, "tests", "fortran", "test_mopac_hip.F90")
FORTRAN_TEST_EXE = os.path.join(ROOT, "tests", "fortran", "_build", "test_mopac_hip")
ARCH = os.environ.get("MHIP_ARCH", "gfx950")

SOURCES = [
    "k_dense.hip",
    "k_fock.hip",
    "k_eigen.hip",
    "k_pseudo.hip",
    "k_occvirt.hip",
    "k_split.hip",
    "k_mlp.hip",
    "k_batch.hip",
    "k_guess.hip",
    "mozyme_scf.hip",
    "k_mozyme.hip",
    "scf_driver.hip",
    "dropin.hip",
    "noalloc.hip",
    "pm7_hamiltonian.cpp",
    "pm7_corrections.cpp",
    "den_io.cpp",
]

# The MOZYME kernels reproduce the reference's list-building, cutoff and
# rotation arithmetic bit for bit (a last-bit change in a rotation's alpha moves
# its beta far more through 1 - alpha^2), so their file is compiled without
# floating-point contraction, like the reference's x86-64 build.
FILE_FLAGS = {"k_mozyme.hip": ["-ffp-contract=off"]}

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise FileNotFoundError(f"libmopachip source missing: {path}")
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + FLAGS + FILE_FLAGS.get(src, []) + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + FLAGS + ["-x", "c++", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libmopachip.so failed")
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> str:
    """The Fortran host module (ISO_C_BINDING interfaces + reference-surface
    wrappers) and the Fortran test program linked against libmopachip.so.
    Skipped when no Fortran compiler is installed."""
    if not os.path.exists(FC):
        return ""
    fdir = os.path.join(HERE, "_build", "fortran")
    tdir = os.path.dirname(FORTRAN_TEST_EXE)
    os.makedirs(fdir, exist_ok=True)
    os.makedirs(tdir, exist_ok=True)
    obj = os.path.join(fdir, "mopac_hip.o")
    cmds = [
        [FC, "-O2", "-fPIC", "-module-dir", fdir, "-c", FORTRAN_MODULE, "-o", obj],
        [FC, "-O2", "-I", fdir, FORTRAN_TEST, obj, "-L", HERE, "-lmopachip",
         "-Wl,-rpath,$ORIGIN/../../../mopac_gpu_amd", "-o", FORTRAN_TEST_EXE],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("Fortran host build failed")
    return FORTRAN_TEST_EXE


if __name__ == "__main__":
    print(build(verbose=True))
