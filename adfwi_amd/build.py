"""Build recipe for libadfwi_amd.so (gfx950) -- plain hipcc, no torch extension machinery.

    python -m adfwi_amd.build          # or adfwi_amd.build.build()

The library is written next to this file so that it travels with the source
tree.  Flags:
  -ffp-contract=off   keep the reference's separately rounded multiply/add order
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions (v_pk_*_f32) in generated code
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libadfwi_amd.so")
SOURCES = ["common.hip", "acoustic.hip", "elastic.hip", "records.hip"]
ARCH = os.environ.get("ADFWI_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def device_flags():
    """Code-generation flags shared by the library build and device_asm."""
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def build(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + device_flags() + ["-fPIC", "-shared", "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def device_asm(source: str, out: str) -> str:
    """gfx950 assembly of one source file (with its kernel metadata: registers, spills, LDS),
    compiled with the library's flags."""
    cmd = [hipcc()] + device_flags() + ["--cuda-device-only", "-S", "-o", out, os.path.join(CSRC, source)]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return out


if __name__ == "__main__":
    build()
