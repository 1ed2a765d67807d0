"""Builds the gfx950 library (libradhip.so) and the host check library in place.

Usage: python -m radiacuda_amd.build   (or radiacuda_amd.build.build()).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled (-packed-fp32-ops) on this target.
# Multiply-adds are contracted into FMAs: the integrals are written in
# cancellation-free forms, so contraction changes results only at rounding level.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
                "-ffp-contract=fast", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)


def build(verbose=True):
    objs = []
    for src in ["rh_kernels.hip", "rh_adjoint.hip", "rh_noalloc.hip", "rh_api.cpp", "rh_geom.cpp"]:
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + DEVICE_FLAGS + lang + ["-c", src, "-o", obj])
        objs.append(obj)
    out = os.path.join(HERE, "libradhip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    chk = os.path.join(HERE, "_rhcheck.so")
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o",
          chk + ".tmp", "rh_hostcheck.cpp", "rh_geom.cpp"])
    os.replace(chk + ".tmp", chk)
    return out


if __name__ == "__main__":
    build()
