"""Build the gfx950 library libgamhip.so in-tree with hipcc.

The library is compiled from mage_amd/csrc/*.hip for --offload-arch=gfx950
and placed next to this file, so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgamhip.so")
SOURCES = ["gam_api.hip", "gam_step.hip", "gam_grid.hip", "gam_adj.hip", "gam_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib(verbose=False):
    """Compile every .hip source and link libgamhip.so (always rebuilds)."""
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for s in SOURCES:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


# HDF5 C library used by the restart library (the image ships it under
# /opt/conda); without it libgamres is not built and restarts are unavailable.
HDF5_PREFIX = os.environ.get("GAM_HDF5_PREFIX", "/opt/conda")
RES_LIB = os.path.join(HERE, "libgamres.so")


def build_restart_lib():
    """Compile libgamres.so (restart files over HDF5) if HDF5 is present;
    returns its path or None."""
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5_hl.h")):
        return None
    tmp = RES_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + inc,
          os.path.join(CSRC, "gam_restart.c"), "-o", tmp, "-L" + HERE, "-lgamhip",
          "-L" + lib, "-lhdf5_hl", "-lhdf5", "-Wl,-rpath," + "$ORIGIN", "-Wl,-rpath," + lib])
    os.replace(tmp, RES_LIB)
    return RES_LIB


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
    print(build_restart_lib())
