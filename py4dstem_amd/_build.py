"""In-tree build of the HIP library ``libpty_hip.so`` for gfx950."""

from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "libpty_hip.so")
SOURCES = [os.path.join(CSRC, "pty_hip.hip")]
HEADERS = [os.path.join(CSRC, "pty_core.h"),
           os.path.join(os.path.dirname(HERE), "include", "pty_hip.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed-FP32 VALU forms are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build libpty_hip.so")


def build(verbose=False, out=LIB_PATH):
    """Compile libpty_hip.so (always recompiles)."""
    cmd = [hipcc_path(), *HIPCC_FLAGS, "-o", out + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
