"""Build recipe for the Fortran layer: the ISO_C_BINDING module (adz_hip.F90), the drop-in
module with the reference's adz_* signatures (adz_gpu.F90), its C-callable entry points
(adz_gpu_capi.F90), the shared library of the three (libadz_gpu.so, used from Python by
gem_amd.adz_fortran) and the host driver (adz_f_driver.F90), all linked against
libadz_hip.so.  In-tree outputs: gem_amd/fortran/_build/*.o, *.mod,
gem_amd/fortran/libadz_gpu.so and gem_amd/fortran/adz_f_driver.

    python gem_amd/fortran/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
DRIVER = os.path.join(HERE, "adz_f_driver")
SHLIB = os.path.join(HERE, "libadz_gpu.so")
MODULES = ["adz_hip.F90", "adz_gpu.F90", "adz_gpu_capi.F90"]


def compiler():
    for c in ("amdflang", "/opt/rocm/bin/amdflang", "flang"):
        p = shutil.which(c)
        if p:
            return p
    return None


def build(verbose=False):
    fc = compiler()
    if fc is None:
        print("fortran: no amdflang/flang on PATH; Fortran layer not built", flush=True)
        return None
    os.makedirs(OUT, exist_ok=True)
    flags = ["-O2", "-fPIC", "-module-dir", OUT]
    objs = []
    for src in MODULES + ["adz_f_driver.F90"]:
        obj = os.path.join(OUT, src.replace(".F90", ".o"))
        cmd = [fc] + flags + ["-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    libs = ["-L" + PKG, "-ladz_hip", "-Wl,-rpath,$ORIGIN/..", "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
            "-Wl,-rpath," + os.path.join(rocm, "lib")]
    shared = [fc, "-shared", "-o", SHLIB + ".tmp"] + objs[:len(MODULES)] + libs
    link = [fc, "-o", DRIVER + ".tmp"] + objs + libs
    for cmd, out in ((shared, SHLIB), (link, DRIVER)):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
