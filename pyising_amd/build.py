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


OBJ_DIR = HERE.parent / "_build"

FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # the device code's LDS atomics are issued by one lane with a wave-uniform address
         # (queue appends); the atomic optimizer's per-lane scan around them is pure overhead
         "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]


def command(out: Path = OUT) -> list[str]:
    """The one-step form of the build (every source into the library), for reference; build()
    compiles the sources in parallel with the same flags and links them."""
    return [hipcc(), *FLAGS, "-shared", f"-I{CSRC}", f"-I{INCLUDE}", "-o", str(out)] + \
        [str(s) for s in sources()]


def _run(cmd: list[str]) -> str:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    return res.stdout + res.stderr


def build(verbose: bool = False) -> Path:
    """Compile every source (always; a build is never skipped), one hipcc per source in
    parallel, link them and replace the library."""
    from concurrent.futures import ThreadPoolExecutor

    OBJ_DIR.mkdir(exist_ok=True)
    srcs = sources()
    objs = [OBJ_DIR / (s.stem + ".o") for s in srcs]
    cmds = [[hipcc(), *FLAGS, f"-I{CSRC}", f"-I{INCLUDE}", "-c", "-o", str(o), str(s)]
            for s, o in zip(srcs, objs)]
    workers = max(1, min(len(cmds), os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as pool:
        logs = list(pool.map(_run, cmds))
    tmp = OUT.with_suffix(".so.tmp")
    logs.append(_run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", str(tmp)] +
                     [str(o) for o in objs]))
    if verbose:
        noise = "packed-fp32-ops"
        msg = "\n".join(ln for log in logs for ln in log.splitlines() if noise not in ln)
        if msg.strip():
            sys.stderr.write(msg + "\n")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
