"""Build of libnttm.so (hipcc, gfx950) and of the oracle (make).

``build_lib()`` compiles every HIP source under ``ntt_amd/csrc`` into ``ntt_amd/libnttm.so`` in one hipcc call.
Packed FP32 instructions are disabled (``-packed-fp32-ops``) and FP contraction is off, so the FP64 residue arithmetic
is exactly the sequence written in ``csrc/arith.h``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libnttm.so")
SOURCES = ["api.hip", "ntt.hip", "crt.hip", "epipe.hip", "digits.hip", "mfma.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def build_lib(verbose=False, out=None, defines=()):
    """compile libnttm.so (or `out`, with extra -D defines, for a comparison build)"""
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    target = out or LIB
    tmp = target + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, target)
    return target


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


if __name__ == "__main__":
    build_lib(verbose=True)
    build_oracle()
    print("built", LIB)
