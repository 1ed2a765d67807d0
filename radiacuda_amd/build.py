"""Builds the gfx950 library (libradhip.so) and the host check library in place.

Usage: python -m radiacuda_amd.build   (or radiacuda_amd.build.build()).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: hand-free of v_pk_{add,mul,fma}_f32 on this target.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
                "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)


def build(verbose=True):
    objs = []
    for src in ["rh_kernels.hip", "rh_api.cpp", "rh_geom.cpp"]:
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
