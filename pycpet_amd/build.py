"""Builds libcpet_hip.so (gfx950) in-tree with hipcc.

    python -m pycpet_amd.build          # or build() from __graft_entry__

Each .hip translation unit is compiled with hipcc and each host-only .cpp
translation unit (text I/O) with g++, in parallel, and the objects are linked
into pycpet_amd/libcpet_hip.so.  Packed FP32 instructions are disabled for
the device pass (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB_NAME = "libcpet_hip.so"
ARCH = os.environ.get("CPET_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function"]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return exe


def _compile(src: Path, obj: Path) -> None:
    if src.suffix == ".cpp":
        cxx = shutil.which("g++") or "g++"
        cmd = [cxx, *CXX_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [_hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed for {src.name}:\n{res.stderr}")
    # surface real warnings, but not the host pass's unknown-feature notice
    msgs = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")


def build(verbose: bool = True) -> Path:
    """Compiles every csrc/*.hip and csrc/*.cpp and links pycpet_amd/libcpet_hip.so."""
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    out = PKG / LIB_NAME
    tmp = PKG / (LIB_NAME + ".tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, out)
    for o in objs:
        o.unlink(missing_ok=True)
    try:
        objdir.rmdir()
    except OSError:
        pass
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
