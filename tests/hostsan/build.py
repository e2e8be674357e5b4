"""Build the host-sanitizer harness: libgbg's host code compiled with
AddressSanitizer (device code left out: --offload-host-only), linked against
the CPU stand-in runtime hip_stub.cpp, and the driver drive.cpp.  Output in
tests/hostsan/_build.  Test infrastructure only; every hipcc line puts
-fsanitize= behind -Xarch_host (host code only)."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "gbgperiodicfields.jl_amd", "csrc")
INC = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
CLANGXX = os.path.join(ROCM, "llvm", "bin", "clang++")
SAN = ["-fsanitize=address", "-fno-omit-frame-pointer"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r.stdout


def compile_unit(src):
    obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
    run([HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "--offload-host-only",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Xarch_host", SAN[0], "-Xarch_host", SAN[1], "-I", INC, "-c", src, "-o", obj])
    return obj


def build() -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_unit, srcs))
    # a host-only unit still names its (absent) device image; give each name a
    # placeholder the stand-in runtime's registration ignores
    names = set()
    for o in objs:
        for line in run(["nm", "-u", o]).splitlines():
            sym = line.split()[-1]
            if sym.startswith("__hip_fatbin"):
                names.add(sym)
    fat_c = os.path.join(OUT, "fatbin_placeholders.c")
    with open(fat_c, "w") as fh:
        for n in sorted(names):
            fh.write(f"const char {n}[16] = {{0}};\n")
    fat_o = os.path.join(OUT, "fatbin_placeholders.o")
    run([CLANGXX, "-x", "c", "-fPIC", "-c", fat_c, "-o", fat_o])
    stub = os.path.join(OUT, "libhip_stub.so")
    run([CLANGXX, "-std=c++17", "-O1", "-g", "-fPIC", "-shared", *SAN, "-D__HIP_PLATFORM_AMD__",
         "-I", os.path.join(ROCM, "include"), os.path.join(HERE, "hip_stub.cpp"), "-o", stub])
    lib = os.path.join(OUT, "libgbg_hostsan.so")
    run([CLANGXX, "-shared", "-fPIC", *SAN, *objs, fat_o, "-L", OUT, "-lhip_stub",
         "-Wl,-rpath,$ORIGIN", "-o", lib])
    drv = os.path.join(OUT, "drive")
    run([CLANGXX, "-std=c++17", "-O1", "-g", *SAN, "-D__HIP_PLATFORM_AMD__", "-I", INC, "-I",
         os.path.join(ROCM, "include"), os.path.join(HERE, "drive.cpp"), "-L", OUT,
         "-lgbg_hostsan", "-lhip_stub", "-Wl,-rpath,$ORIGIN", "-o", drv])
    return drv


if __name__ == "__main__":
    print(build())
    sys.exit(0)
