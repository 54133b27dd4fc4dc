"""Build ``libising_hip.so`` in-tree with hipcc for gfx950 (``python -m pyising_amd.build``).

One shared library from every ``csrc/*.hip`` file; the C ABI it exports is declared in
``include/ising_hip.h``. Packed-FP32 instruction selection is disabled for all device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "libising_hip.so"
ARCH = os.environ.get("ISING_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def command(out: Path = OUT) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Wall", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # the device code's LDS atomics are issued by one lane with a wave-uniform address
            # (queue appends); the atomic optimizer's per-lane scan around them is pure overhead
            "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
            f"-I{CSRC}", f"-I{INCLUDE}", "-o", str(out)] + [str(s) for s in sources()]


def build(verbose: bool = False) -> Path:
    """Compile every source (always; a build is never skipped) and replace the library."""
    tmp = OUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    if verbose:
        noise = "packed-fp32-ops"
        msg = "\n".join(ln for ln in (res.stdout + res.stderr).splitlines() if noise not in ln)
        if msg.strip():
            sys.stderr.write(msg + "\n")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
