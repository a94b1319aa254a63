"""Builds libgct_amd.so in-tree for gfx950 with hipcc / amdclang++.

    python -m gct_amd.build            # library only
    python -m gct_amd.build --all      # library, C++ API test programs

The library is written to gct_amd/libgct_amd.so; objects go to gct_amd/_build/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "gct_amd"
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libgct_amd.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("GCT_OFFLOAD_ARCH", "gfx950")

HIPCC = str(ROCM / "bin" / "hipcc")
CLANGXX = str(ROCM / "llvm" / "bin" / "clang++")

COMMON = ["-O3", "-std=c++20", "-fPIC", f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
# Packed FP32 instructions are disabled in device code.  Multiply-adds are
# fused only within an expression (-ffp-contract=on): fusing across
# statements would let the compiler round a kernel's arithmetic differently
# in each template instance, and the backprojection variants must agree bit
# for bit; kernels spell out the fmaf they rely on.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=on", "-munsafe-fp-atomics"]
HOST = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def _compile(src: Path) -> Path:
    obj = OUT / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CLANGXX, *COMMON, *HOST, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs),
          f"-L{ROCM / 'lib'}", "-lamdhip64", "-lpthread", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests() -> list[Path]:
    """C++ programs that exercise the header-only GCT:: API over the C ABI."""
    src_dir = ROOT / "tests" / "cpp"
    out_dir = src_dir / "_build"
    out_dir.mkdir(parents=True, exist_ok=True)
    built = []
    for src in sorted(src_dir.glob("*.cpp")):
        exe = out_dir / src.stem
        _run(["g++", "-O2", "-std=c++20", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
              f"-L{PKG}", "-lgct_amd", f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{ROCM / 'lib'}"])
        built.append(exe)
    return built


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) and, where the reference tree is
    present, oracle/_ref/libgctref.so from the reference's own sources."""
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", str(ROOT / "oracle")])
    if Path("/root/reference/src/cpp/recon/ramp_filter.cpp").exists():
        _run([make, "-s", "-C", str(ROOT / "oracle"), "ref"])


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build_library(a.j))
    if a.all:
        build_oracle()
        for e in build_cpp_tests():
            print(e)


if __name__ == "__main__":
    main()
