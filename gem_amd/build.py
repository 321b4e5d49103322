"""Build recipe for libadz_hip.so (gfx950) -- in-tree, so the library travels with the
repository snapshot.

    python -m gem_amd.build            # compile gem_amd/libadz_hip.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libadz_hip.so")
SOURCES = ["adz_interp.hip", "adz_traject.hip", "adz_batched.hip", "adz_bc.hip", "adz_flux.hip", "adz_qv.hip", "adz_ilmc.hip", "adz_airmass.hip"]

# -ffp-contract=off: device FMAs are written explicitly, so operation order is fixed.
# -packed-fp32-ops disabled: no v_pk_*_f32 instructions are generated.
HIPFLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-pass-failed",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False):
    cmd = ["hipcc"] + HIPFLAGS + ["-o", LIB + ".tmp"] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libadz_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
