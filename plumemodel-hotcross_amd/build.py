#!/usr/bin/env python3
"""Builds libhotcross.so (the HIP kernels and the C-ABI of include/hotcross.h)
for gfx950 in place, with hipcc:

  hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off
        -Xclang -target-feature -Xclang -packed-fp32-ops csrc/*.hip

-ffp-contract=off keeps the reference's rounding (no fused multiply-adds are
formed from separate products and sums); packed FP32 instructions are disabled
for all device code.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhotcross.so")


def build(verbose=False):
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-Wno-unused-value", "-o", LIB + ".tmp"] + srcs
    r = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l or "error" in l)
    if verbose or r.returncode:
        sys.stdout.write(out + "\n")
    if r.returncode:
        raise RuntimeError("hipcc failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_fortran():
    """The Fortran host layer (ISO_C_BINDING interfaces, drop-in subroutines
    with the reference's names, and the driver) with amdflang, linked against
    libhotcross.so."""
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for f in ("hotcross_capi.f90", "hotcross_dropin.f90", "hotcross_driver.f90"):
        o = os.path.join(bdir, f.replace(".f90", ".o"))
        subprocess.run(["amdflang", "-O2", "-J" + bdir, "-c", os.path.join(fdir, f), "-o", o], check=True)
        objs.append(o)
    exe = os.path.join(fdir, "hotcross_driver")
    subprocess.run(["amdflang", "-o", exe] + objs[::-1] + ["-L" + HERE, "-lhotcross",
                                                          "-Wl,-rpath,$ORIGIN/.."], check=True)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
