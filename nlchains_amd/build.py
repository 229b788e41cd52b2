"""Build recipe for libnlchains_amd.so (gfx950) and the nlchains CLI.

Run as ``python -m nlchains_amd.build`` or through ``__graft_entry__.build()``.
Every HIP translation unit is compiled in parallel with hipcc into
``nlchains_amd/_build/`` and linked into ``nlchains_amd/libnlchains_amd.so``
(in-tree, so it travels with the source tree).  The CLI ``nlchains_amd/nlchains``
links against that library.
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
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libnlchains_amd.so"
CLI = PKG / "nlchains"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def _sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    cmds = []
    for src in srcs:
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        if src.suffix == ".hip":
            cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        else:  # host-only C++ that uses the HIP runtime API
            cmds.append([HIPCC, *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs), "-lpthread", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> Path | None:
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return None
    tmp = CLI.with_name(CLI.name + ".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", *map(str, srcs), "-o", str(tmp),
          f"-L{PKG}", "-lnlchains_amd", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle() -> None:
    _run(["sh", str(ROOT / "oracle" / "build.sh")])


def build_all() -> None:
    build_oracle()
    build_library()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_library, "cli": build_cli, "oracle": build_oracle}[w]()
    print("built:", LIB, "(exists)" if LIB.exists() else "(missing)")
