"""Build the gfx950 library in-tree: strata_amd/libstrata_hip.so.

One hipcc invocation over the HIP sources in strata_amd/csrc. Flags:
  -ffp-contract=off   no fused multiply-add: every float32 operation rounds
                      once, in the reference's order
  packed-fp32-ops off the device pass must not emit v_pk_{add,mul,fma}_f32
  amdgpu-mfma-vgpr-form
                      matrix-core accumulators in ordinary VGPRs: with the
                      AGPR form the MFMA Correlate's loop renamed its four
                      accumulators every weight row and moved them back
                      (about 30 accumulator moves a row)
  (no fast-math, denormals kept, correctly rounded f32 divide and sqrt: the
  HIP defaults)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstrata_hip.so")
SOURCES = ["kernels.hip", "runtime.hip", "batch.hip", "serve.hip", "reduce.hip", "rawio.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile the library (always) and return its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only feature as unknown; that
    # line is expected and dropped from the log.
    log = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{log}")
    if verbose and log.strip():
        print(log)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
