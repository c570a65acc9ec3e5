#!/usr/bin/env python3
"""Build recipe for oracle/_ref: the reference's own Fortran, compiled in place.

TEST INFRASTRUCTURE.  For each grid size it creates oracle/_ref/g<nx>x<ny>x<nz>/
holding symbolic links to the reference source files under /root/reference
(nothing is copied), the reference's compile-time grid configuration
array_size.inc (array_size.inc:1-6 fixes nx, ny, nz as parameters; the recipe
writes the one line that selects the grid, as the reference's own commented
alternatives at array_size.inc:4-6 do), and compiles with amdflang:

  amdflang -c -O3 -ffixed-form -ffixed-line-length=132 -fdefault-real-8
           -fopenmp -fd-lines-as-comments -fPIC -mcmodel=medium

(-fdefault-real-8 is flang's spelling of PGI's -r8, Hotcross-SAPELO2.sh:15;
-mcmodel=medium because the module and main-program arrays are static and
pass 2 GiB at 256^3).

hotcross80-revised2014.f holds FORCES, PRESSURE, MOMENTUM_U/V/W, REFRESH and
INITIALIZE together with `program hotcross`, whose body includes the run's case
file (case.inc:1 -> DanteExp1.case) and its output include write.inc
(hotcross80-revised2014.f:118,257,354,361).  Neither file is in the reference;
both are read only inside `program hotcross`, which the harness never runs,
and the recipe writes them empty.  The netCDF writers are not built.

smagorinsky_split_mixing-revised2014.f (VISCOSITY3D) applies the COMPLEX
intrinsic csqrt to a REAL argument at :112, which PGI accepts by promoting the
argument to COMPLEX and flang rejects.  The file is compiled in place with the
preprocessor on and one definition,

  -cpp -D'csqrt(x)=sqrt((x)+(0d0,0d0))'

which is that promotion written out: the generic sqrt of a COMPLEX(8) is the
principal root, and the assignment to the REAL ss_v keeps its real part, as
under PGI -r8 (whose precision for the promoted csqrt is not documented).

The library is loaded with RTLD_LAZY so the unresolved writer symbols
referenced by `program hotcross` are never bound.

usage: python oracle/build_ref.py 16 12 10 [nx ny nz ...]
"""
import os
import subprocess
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

SOURCES = [
    "modules80-revised2014.f",
    "hotcross80-revised2014.f",
    "source80.f",
    "potential_density80.f",
    "initialize_profiles80.f",
    "hs3crt_composite80.f",
    "mixing80-revised2014.f",
    "body_force80.f",
    "sponge80.f",
    "mpdata-hotcross80.f",
    "smagorinsky_split_mixing-revised2014.f",
]
# per-file additions to FFLAGS (see the module docstring)
EXTRA = {"smagorinsky_split_mixing-revised2014.f": ["-cpp", "-Dcsqrt(x)=sqrt((x)+(0d0,0d0))"]}
INCLUDES = ["case.inc", "site_forcing.inc", "source80.inc"]
FFLAGS = ["-O3", "-ffixed-form", "-ffixed-line-length=132", "-fdefault-real-8",
          "-fopenmp", "-fd-lines-as-comments", "-fPIC", "-mcmodel=medium"]


def libpath(nx, ny, nz):
    return os.path.join(OUT, f"g{nx}x{ny}x{nz}", "libhcref.so")


def build(nx, ny, nz, quiet=True):
    if not os.path.isdir(REF):
        raise RuntimeError("reference sources not present")
    d = os.path.join(OUT, f"g{nx}x{ny}x{nz}")
    os.makedirs(d, exist_ok=True)
    for f in SOURCES + INCLUDES:
        link = os.path.join(d, f)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(os.path.join(REF, f), link)
    with open(os.path.join(d, "array_size.inc"), "w") as fh:
        fh.write("      integer nx, ny, nz\n")
        fh.write(f"         parameter ( nx = {nx}, ny = {ny}, nz = {nz})\n")
    for f in ("DanteExp1.case", "write.inc"):
        open(os.path.join(d, f), "w").close()
    objs = []
    run = lambda cmd: subprocess.run(cmd, cwd=d, check=True,
                                     stdout=subprocess.PIPE if quiet else None,
                                     stderr=subprocess.STDOUT if quiet else None)
    for f in SOURCES:
        run(["amdflang", "-c"] + FFLAGS + EXTRA.get(f, []) + [f])
        objs.append(f.replace(".f", ".o"))
    run(["amdflang", "-c", "-O2", "-fdefault-real-8", "-fopenmp", "-fPIC",
         "-mcmodel=medium", "-I.", os.path.join(HERE, "ref_harness.f90"), "-o", "ref_harness.o"])
    run(["amdflang", "-shared", "-fopenmp", "-o", "libhcref.so",
         "ref_harness.o"] + objs)
    return libpath(nx, ny, nz)


if __name__ == "__main__":
    a = [int(x) for x in sys.argv[1:]]
    if not a or len(a) % 3:
        sys.exit(__doc__)
    for q in range(0, len(a), 3):
        print(build(*a[q:q + 3], quiet=False))
