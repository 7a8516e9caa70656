"""Build libxsmd.so (gfx950 HIP kernels + C ABI) in-tree.

    python -m exastamp_amd.build          # compile every source, link the library

The library lands at exastamp_amd/libxsmd.so; objects go to exastamp_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libxsmd.so"

SOURCES = ["capi.hip", "k_neighbors.hip", "k_integrate.hip", "k_pair.hip", "k_adjoint.hip", "k_block.hip", "setfl.cpp"]

ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code (-packed-fp32-ops)
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path, obj: Path) -> tuple[str, int, str]:
    cmd = [hipcc(), "-x", "hip", "-c", str(src), "-o", str(obj), "-I", str(ROOT / "include"),
           "-I", str(CSRC)] + HIP_FLAGS
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src.name, p.returncode, (p.stdout + p.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, rc, out in ex.map(lambda so: _compile(*so), zip(srcs, objs)):
            if rc != 0:
                errors.append(f"--- {name} ---\n{out}")
            elif verbose and out.strip():
                # warnings about the host pass not knowing packed-fp32-ops are expected
                lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
                if lines:
                    print(f"--- {name} ---\n" + "\n".join(lines))
    if errors:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(errors))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
