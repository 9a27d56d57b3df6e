"""In-tree build of the native library ``dlux_amd/_native/libdlux_amd.so``.

Every ``csrc/*.hip`` file is compiled for gfx950 with hipcc (in parallel) and
linked into one shared library whose ``extern "C"`` symbols are declared in
``include/dlux_amd.h``.  No CUDA, no hipify, one target architecture.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
NATIVE = PKG / "_native"
LIB = NATIVE / "libdlux_amd.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for all device code.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: Path, obj: Path, extra: list[str]) -> tuple[Path, str]:
    cmd = [HIPCC, *DEVICE_FLAGS, *extra, "-I", str(ROOT / "include"), "-c", str(src),
           "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{p.stderr}")
    # hipcc's host pass warns that it does not know the device-only feature.
    noise = "is not a recognized feature for this target"
    msgs = "\n".join(l for l in p.stderr.splitlines() if noise not in l)
    return obj, msgs


def build_native(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile every HIP source and link ``libdlux_amd.so``; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    NATIVE.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    extra = list(extra or [])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        futs = [ex.submit(_compile, s, BUILD / (s.stem + ".o"), extra) for s in srcs]
        objs = []
        for f in futs:
            obj, msgs = f.result()
            objs.append(obj)
            if verbose and msgs.strip():
                print(msgs)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{p.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build_native(verbose=True)
