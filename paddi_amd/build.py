"""Build step of the HIP library (hipcc, gfx950) and of the CPU oracle (gcc).

``python -m paddi_amd.build`` compiles in place:
  paddi_amd/libpaddi_hip.so   kernels.hip + engine.cpp (+ the learned-model kernels)
  oracle/liboracle.so         the test-only CPU oracle (make -C oracle)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpaddi_hip.so")

HIP_SOURCES = ["kernels.hip", "mlp.hip", "scratch.hip", "sax.hip", "engine.cpp",
               "learned.cpp"]  # the library
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
    # packed FP32 VALU instructions (v_pk_*_f32) are disabled: some of their forms have given
    # silently wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.sep not in p or os.path.exists(p)):
            return p
    raise RuntimeError("hipcc not found")


def build_hip(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "hip-link" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


CLI = os.path.join(HERE, "paddi_main")


def build_cli():
    """The command-line drop-in (csrc/paddi_main.cpp), linked against the library."""
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "paddi_main.cpp"),
           "-L" + HERE, "-lpaddi_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed:\n{' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-C", d, "-B"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}{r.stderr}")


def build_all(verbose=False):
    build_hip(verbose)
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
