"""Compile MOPAC's own SCF routines from the reference sources into
oracle/_ref/libmopac_ref.so — TEST INFRASTRUCTURE (a checker for the C
restatement, never part of the product).

    python oracle/build_ref.py [--src /root/reference/src]

Only the routines on the SCF inner loop and the small modules they read are
compiled, with amdflang, straight from where they lie.  The reference's own
CMake build is not used.  Symbols those files reference but never reach on
this path stay unresolved (the COSMO Fock term addfck, guarded by
useps = .false.; ijbo, used only when lijbo is false; mopend, reada and timer,
reached only on error or by keywords the tests do not set; the d-orbital and
solid-state branches of the integral and correction files), so the library is
opened lazily (RTLD_LAZY); the BLAS/LAPACK routines pulay calls resolve from
MKL, which the oracle loader makes global.
"""
import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
SO = os.path.join(OUT, "libmopac_ref.so")
FC = os.environ.get("MOPAC_REF_FC", "/opt/rocm/bin/amdflang")

# order matters: modules before their users
SOURCES = [
    "molkst_C.F90",
    "solvation/cosmo_C.F90",
    "chanel_C.F90",
    "SCF/iter_C.F90",
    "SCF/fock1.F90",
    "integrals/jab.F90",
    "integrals/kab.F90",
    "SCF/fock2.F90",
    "integrals/helect.F90",
    "SCF/cnvg.F90",
    "SCF/pulay.F90",
    "matrix/diag.F90",
    "matrix/osinv.F90",
    # MOZYME kernels and the modules they read
    "Common_arrays_C.F90",
    "models/parameters_C.F90",
    "MOZYME/MOZYME_C.F90",
    "MOZYME/epseta.F90",
    "MOZYME/diagg1.F90",
    "MOZYME/diagg2.F90",
    "MOZYME/density_for_MOZYME.F90",
    "overlaps_C.F90",
    "utilities/ijbo.F90",
    "MOZYME/helecz.F90",
    "MOZYME/cnvgz.F90",
    # the sparse Fock build and what it reads (linear_cosmo only for the
    # module interface of its COSMO term, never reached with useps false)
    "funcon_C.F90",
    "solvation/afmm_C.F90",
    "solvation/linear_cosmo.F90",
    "MOZYME/chrge_for_MOZYME.F90",
    "MOZYME/jab_for_MOZYME.F90",
    "MOZYME/kab_for_MOZYME.F90",
    "MOZYME/fock1_for_MOZYME.F90",
    "integrals/mndod_C.F90",
    "integrals/mndod.F90",
    "MOZYME/fock2z.F90",
    "MOZYME/tidy.F90",
    "MOZYME/check.F90",
    "MOZYME/eimp.F90",
    # the PM7 producer's reference: parameters, derived multipoles, pair
    # integrals (rotate/rotatd/reppd/spcore/ccrep), overlaps (h1elec/diat/set)
    "conref_C.F90",
    "INDO/reimers_C.F90",
    "elemts_C.F90",
    "chemistry/atomradii_C.F90",
    "models/parameters_for_PM7_C.F90",
    "integrals/sp_two_electron.F90",
    "models/calpar.F90",
    "integrals/rotate.F90",
    "integrals/h1elec.F90",
    "integrals/diat.F90",
    "integrals/set.F90",
    "integrals/coe.F90",
    "integrals/ccrep.F90",
    "input/digit.F90",
    "utilities/reada.F90",
    "integrals/solrot.F90",
    "MOZYME/outer1.F90",
    "MOZYME/outer2.F90",
    "MOZYME/fillij.F90",
    "MOZYME/add_more_interactions.F90",
    "MOZYME/hcore_for_MOZYME.F90",
    # PM7's post-SCF corrections (dispersion, hydrogen bonds) and the bonding
    # topology they read
    "chemistry/geochk_bits_1.F90",
    "geometry/bangle.F90",
    "geometry/dihed.F90",
    "geometry/dang.F90",
    "corrections/set_up_dentate.F90",
    "corrections/H_bond_correction_PM6_DH_Dispersion.F90",
    "corrections/H_bond_correction_bits.F90",
    "corrections/H_bond_correction_EH_plus.F90",
    "corrections/Hydrogen_bond_corrections.F90",
]
HARNESSES = ["ref_harness.F90", "ref_harness_mozyme.F90", "ref_harness_pm7.F90"]


def build(src="/root/reference/src"):
    if not os.path.isdir(src):
        raise FileNotFoundError(f"reference sources not found at {src}")
    mod = os.path.join(OUT, "mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    flags = ["-O2", "-fPIC", "-module-dir", mod, "-I", mod]
    for rel in SOURCES + [os.path.join(HERE, h) for h in HARNESSES]:
        path = rel if os.path.isabs(rel) else os.path.join(src, rel)
        obj = os.path.join(OUT, os.path.basename(path).replace(".F90", ".o"))
        subprocess.run([FC] + flags + ["-c", path, "-o", obj], check=True, cwd=OUT)
        objs.append(obj)
    subprocess.run([FC, "-shared", "-o", SO] + objs, check=True, cwd=OUT)
    return SO


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--src", default="/root/reference/src")
    a = ap.parse_args()
    print(build(a.src))
    sys.exit(0)
