"""In-tree build of the engine libraries for gfx950.

Produces, under totem_amd/lib/:
  libtotem.so                  -- kernels + host engine + C ABI (include/totem_abi.h)
  libwaveCudalobato.so         -- myProjectfun -> totem_project_lobato
  libwaveCudaPengAndCorr.so    -- myProjectfun -> totem_project_peng
  libwaveCudalobatoP.so        -- myProjectfun -> totem_potential_lobato
  libwaveCudaPengAndCorrP.so   -- myProjectfun -> totem_potential_peng
(the four names of the reference's mex products, mex_cuda.m:19-38).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # never emit packed FP32 arithmetic (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

VARIANTS = {
    "libwaveCudalobato.so": "totem_project_lobato",
    "libwaveCudaPengAndCorr.so": "totem_project_peng",
    "libwaveCudalobatoP.so": "totem_potential_lobato",
    "libwaveCudaPengAndCorrP.so": "totem_potential_peng",
}


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in ("kernels.hip", "engine.hip", "server.hip", "hrtem.hip", "specimen.hip"):
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(SRC, src), "-o", obj])
        objs.append(obj)
    core = os.path.join(LIB, "libtotem.so")
    tmp = core + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp])
    os.replace(tmp, core)
    for name, fn in VARIANTS.items():
        out = os.path.join(LIB, name)
        _run(["g++", "-O2", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
              "-DTOTEM_VARIANT=" + fn, os.path.join(SRC, "variant.cpp"), "-L" + LIB, "-ltotem",
              "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return core


if __name__ == "__main__":
    build()
    sys.exit(0)
