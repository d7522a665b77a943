"""Compiles the HIP library nmc_amd/libnmc_hip.so for gfx950 with hipcc.

Flags that affect results:
  -ffp-contract=off   no implicit fma: float expressions round exactly as
                      written, on the host and the device alike;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 ops;
  -munsafe-fp-atomics hardware FP32/FP64 atomic add for the tallies.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnmc_hip.so")
SOURCES = ["nmc_abi.hip", "mcxyz.hip", "mcsub.hip", "layered.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NMC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


# measurement kernels used by bench.py only (atomic-rate ceiling, copy calibration)
PROBE_LIB = os.path.join(HERE, "libnmc_probe.so")
PROBE_SOURCES = ["probe.hip"]


def _hipcc(out: str, sources: list[str], verbose: bool) -> None:
    cmd = [HIPCC, *FLAGS, "-o", out + ".tmp", *[os.path.join(CSRC, s) for s in sources]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(out)}")
    os.replace(out + ".tmp", out)


def build(verbose: bool = False) -> str:
    """Builds both libraries in-tree (always recompiles); returns the main one."""
    _hipcc(LIB, SOURCES, verbose)
    _hipcc(PROBE_LIB, PROBE_SOURCES, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
